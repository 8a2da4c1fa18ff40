"""msc_scc_run_noalloc (the entry point that allocates nothing: its workspace is the
library's static per-device scratch array) against msc_scc_run on the same inputs: two
calls in a row, the second on a smaller graph, so the second call runs over what the first
left in the array.  Labels must be bit-exact, raw and canonical."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(msc, ctx, g, labels):
    dg = msc.DeviceGraph(ctx, g)
    out = msc.DeviceLabels(ctx, g.n)
    try:
        want = msc.run_scc(dg, g.max_deg_vert, g.avg_degree, ctx=ctx, labels=labels)
        got, st = msc.run_scc_noalloc(dg, g.max_deg_vert, g.avg_degree, ctx=ctx, out=out, labels=labels,
                                      stats=True)
        return want, got.to_host(), st
    finally:
        out.free()
        dg.free()


@pytest.mark.parametrize("labels", ["raw", "canonical"])
def test_noalloc_matches_scc_run_twice(msc, gpu_ctx, labels):
    mode = msc.LABELS_RAW if labels == "raw" else msc.LABELS_CANONICAL
    big = msc.gen_rmat(16, 16, seed=5)
    small, _ = msc.gen_planted(5000, 1.5, 400, 1.0, 20000, seed=9)
    print(f"noalloc {labels}: R-MAT s16 n={big.n} m={big.m}, then planted n={small.n} m={small.m}", flush=True)
    for g in (big, small):
        want, got, st = _pair(msc, gpu_ctx, g, mode)
        assert np.array_equal(want, got), f"{int((want != got).sum())} labels differ (n={g.n})"
        assert st.pivot == g.max_deg_vert


def test_noalloc_device_pivot_and_learned(msc, gpu_ctx):
    g = msc.gen_rmat(14, 8, seed=11)
    msc.load_policy(gpu_ctx)
    dg = msc.DeviceGraph(gpu_ctx, g)
    out = msc.DeviceLabels(gpu_ctx, g.n)
    try:
        want = msc.run_scc(dg, ctx=gpu_ctx, policy=msc.POLICY_LEARNED)
        got = msc.run_scc_noalloc(dg, ctx=gpu_ctx, out=out, policy=msc.POLICY_LEARNED).to_host()
        assert np.array_equal(want, got)
    finally:
        out.free()
        dg.free()


def test_noalloc_limits_and_requirements(msc, gpu_ctx):
    # host-resident graph: refused (it would need a staging buffer)
    g = msc.gen_rmat(8, 8, seed=1)
    out = msc.DeviceLabels(gpu_ctx, g.n)
    try:
        with pytest.raises(msc.MscError) as e:
            msc.run_scc_noalloc(g, ctx=gpu_ctx, out=out)
        assert e.value.code == msc.E_INVALID
    finally:
        out.free()
    # more vertices than the scratch array is sized for
    n = (1 << 20) + 1
    big = msc.create_csr(n, [0], [1])
    dg = msc.DeviceGraph(gpu_ctx, big)
    out = msc.DeviceLabels(gpu_ctx, n)
    try:
        with pytest.raises(msc.MscError) as e:
            msc.run_scc_noalloc(dg, ctx=gpu_ctx, out=out)
        assert e.value.code == msc.E_CAPACITY
    finally:
        out.free()
        dg.free()
