"""The autotuner (connectivity_amd/autotune.py): one coordinate-descent pass on a small
graph leaves the labels those of the oracle, and the tuning it picks is one the library
accepts (msc_ctx_set_tuning) and keeps on the context."""
import numpy as np
import pytest


@pytest.mark.gpu
@pytest.mark.parametrize("alg", ["scc", "cc"])
def test_one_pass_keeps_labels(msc, orc, gpu_ctx, alg):
    from connectivity_amd import autotune

    defaults = gpu_ctx.tuning()
    g = msc.gen_rmat(12, 16, seed=21)
    lines = []
    try:
        r = autotune.tune(gpu_ctx, g, alg, msc.POLICY_PLAIN, reps=1, passes=1, log=lines.append)
        assert gpu_ctx.tuning() == r["tuning"]
        assert set(r["tuning"]) == set(autotune.CANDIDATES)
        # every candidate of the pass was timed (each compared its labels with the defaults')
        assert len(r["trials"]) >= 1 + sum(len(v) - 1 for v in autotune.CANDIDATES.values()) - 2
        assert r["ms"] <= r["default_ms"]
        want = orc.canonical(orc.scc(g, vert_cutoff=0) if alg == "scc" else orc.cc(g))
        got = (msc.run_scc if alg == "scc" else msc.run_cc)(g, ctx=gpu_ctx)
        assert np.array_equal(got, want)
    finally:
        gpu_ctx.set_tuning(**defaults)


def test_candidates_are_accepted_values():
    from connectivity_amd import autotune

    # msc_ctx_set_tuning's ranges (include/msc.h): every candidate value lies inside them
    lim = {"blocks_per_cu": (1, 8), "small_push_edges": (0, 1 << 30), "small_cbw_edges": (0, 1 << 30),
           "first_batch": (1, 256), "max_batch": (1, 256), "concurrent_fwbw": (0, 2), "search_batch": (1, 256),
           "local_rounds": (0, 256), "local_us": (1, 100000), "pull_div": (0, 1024),
           "tail_vertices": (0, 65536), "balance": (0, 3), "group_words": (0, 256)}
    for field, values in autotune.CANDIDATES.items():
        lo, hi = lim[field]
        assert all(lo <= v <= hi for v in values), field
