"""Colour tail (msc_tuning.tail_vertices): when FW-BW leaves a small colouring, one
workgroup finds the components among the vertices left from the edges that join them and
labels them, instead of colouring rounds.  Labels must equal the oracle's bit for bit (raw:
trimmed -> itself, pivot's component -> pivot, every other component -> its largest id;
canonical: smallest id), with the tail and with it off, for SCC and WCC.  Inputs whose
joining edges do not fit are declined and coloured in rounds, with the same labels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _random(msc, n, m, seed):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m, dtype=np.int64)
    dst = rng.integers(0, n, m, dtype=np.int64)
    return msc.create_csr(n, src.astype(np.int32), dst.astype(np.int32))


def _graphs(msc):
    return [
        ("rmat12", msc.gen_rmat(12, 8, seed=3)),
        ("rmat16", msc.gen_rmat(16, 16, seed=1)),
        ("rmat18", msc.gen_rmat(18, 16, seed=2)),
        ("random4000", _random(msc, 4000, 6000, 5)),  # many small SCCs left after FW-BW
        ("planted3000", msc.gen_planted(3000, max_size=40, seed=7)[0]),
        ("grid48", msc.gen_grid(48, 48, seed=2)),
    ]


def _run(msc, ctx, kind, g, piv, avg, labels, tail):
    ctx.set_tuning(tail_vertices=65536 if tail else 0)
    run = msc.run_scc if kind == "scc" else msc.run_cc
    return run(g, piv, avg, ctx=ctx, labels=labels, stats=True)


@pytest.mark.parametrize("kind", ["scc", "cc"])
def test_tail_labels_equal_oracle_and_rounds(msc, orc, gpu_ctx, kind):
    saved = gpu_ctx.tuning()
    ran = []
    try:
        for name, g in _graphs(msc):
            piv, avg = orc.pivot_scc(g) if kind == "scc" else orc.pivot_cc(g)
            raw = orc.scc(g, piv, avg, vert_cutoff=0) if kind == "scc" else orc.cc(g, piv, avg)
            for labels, want in ((msc.LABELS_RAW, raw), (msc.LABELS_CANONICAL, orc.canonical(raw))):
                for tail in (True, False):
                    _say(f"case: {kind} {name} labels={labels} tail={tail}")
                    got, st = _run(msc, gpu_ctx, kind, g, piv, avg, labels, tail)
                    bad = np.flatnonzero(got != want)
                    assert bad.size == 0, (name, labels, tail, bad[:5], got[bad[:5]], want[bad[:5]])
                    if not tail:
                        assert st.color_tail == 0
                    elif st.color_tail:
                        ran.append(name)
    finally:
        gpu_ctx.set_tuning(**saved)
    # R-MAT leaves a few sparsely joined vertices; the random digraph's and the grid's give
    # the tail components of several vertices (up to 7 and 28); the planted graph's dense
    # SCCs (~16 000 joining edges) exceed it and are coloured in rounds
    for name in ("rmat12", "rmat16", "rmat18", "random4000", "grid48"):
        assert name in ran, (kind, name, ran)


def test_tail_finds_multi_vertex_components(msc, orc, gpu_ctx):
    """The tail's own components, not only singletons: on the 48 x 48 grid the vertices
    FW-BW leaves hold SCCs of up to 28 vertices, labelled by their largest id."""
    g = msc.gen_grid(48, 48, seed=2)
    piv, avg = orc.pivot_scc(g)
    raw = orc.scc(g, piv, avg, vert_cutoff=0)
    got, st = _run(msc, gpu_ctx, "scc", g, piv, avg, msc.LABELS_RAW, True)
    assert st.color_tail == 1
    assert np.array_equal(got, raw)
    left = raw != piv
    sizes = np.bincount(raw[left])
    assert sizes.max() > 1  # components of several vertices were settled by the tail


def test_tail_declines(msc, orc, gpu_ctx):
    saved = gpu_ctx.tuning()
    try:
        # more vertices than allowed
        g = msc.gen_rmat(16, 16, seed=1)
        piv, avg = orc.pivot_scc(g)
        want = orc.scc(g, piv, avg, vert_cutoff=0)
        gpu_ctx.set_tuning(tail_vertices=8)
        got, st = msc.run_scc(g, piv, avg, ctx=gpu_ctx, labels=msc.LABELS_RAW, stats=True)
        assert st.color_tail == 0 and np.array_equal(got, want)
        # more joining edges than fit (the planted graph's SCCs hold ~10^5 of them)
        gpu_ctx.set_tuning(tail_vertices=65536)
        g = msc.gen_planted(20000, max_size=2000, seed=6)[0]
        for kind in ("scc", "cc"):
            piv, avg = orc.pivot_scc(g) if kind == "scc" else orc.pivot_cc(g)
            raw = orc.scc(g, piv, avg, vert_cutoff=0) if kind == "scc" else orc.cc(g, piv, avg)
            for labels, want in ((msc.LABELS_RAW, raw), (msc.LABELS_CANONICAL, orc.canonical(raw))):
                _say(f"case: decline planted20000 {kind} labels={labels}")
                got, st = _run(msc, gpu_ctx, kind, g, piv, avg, labels, True)
                assert np.array_equal(got, want), (kind, labels)
        # the SCC remainder's joining edges exceed the tail: declined, coloured in rounds
        piv, avg = orc.pivot_scc(g)
        _, st = _run(msc, gpu_ctx, "scc", g, piv, avg, msc.LABELS_RAW, True)
        assert st.color_tail == 0
    finally:
        gpu_ctx.set_tuning(**saved)


def test_tail_repeated_calls_and_device_buffers(msc, orc, gpu_ctx):
    """Back-to-back calls on device-resident buffers (the benchmark's form), alternating
    with a graph the tail declines, keep exact labels: no state leaks between calls."""
    small = msc.gen_rmat(14, 16, seed=4)
    big = msc.gen_planted(20000, max_size=2000, seed=6)[0]
    cases = []
    for g in (small, big):
        piv, avg = orc.pivot_scc(g)
        cases.append((g, piv, avg, orc.canonical(orc.scc(g, piv, avg, vert_cutoff=0)),
                      msc.DeviceGraph(gpu_ctx, g), msc.DeviceLabels(gpu_ctx, g.n)))
    try:
        for it in range(3):
            for g, piv, avg, want, dg, out in cases:
                _say(f"case: repeated call {it} n={g.n}")
                msc.run_scc(dg, piv, avg, ctx=gpu_ctx, out=out)
                assert np.array_equal(out.to_host(), want)
    finally:
        for *_, dg, out in cases:
            out.free()
            dg.free()
