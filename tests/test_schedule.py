"""The plain direction schedule (the reference's ALPHA / BETA switch,
multistep/scc_fwbw.cpp:164-195 forward, :347-378 backward).

CPU: known answers on layered digraphs whose layer sizes sit on the rule's boundaries,
worked out by hand from the reference's text (strict `<`, total n rather than the valid
count, integer n / BETA, the one-way return to top-down).  The reference's own per-level
DEBUG lines cannot be switched on without editing its source (scc_main.cpp:58 defines
DEBUG 0 unconditionally), so these literals pin the oracle's trace.

GPU: with the searches run one after the other (concurrent_fwbw = 0, the reference's
order), the device's level records (phase, level, direction, vertices found) equal the
oracle's trace for FW and BW up to and including the first bottom-up decision.  Later
levels are not comparable: the reference's bottom-up sweep marks vertices while it runs
(scc_fwbw.cpp:118-148), so a sweep can claim more than one layer, depending on order.
"""
import numpy as np
import pytest

from layered import layered

SIZES = [100, 94, 300, 500, 400, 140, 64, 1]

# n = 1600, avg 1.0.  FW (found excludes the root):
#  L1 TD 100: (1600-100)/15 = 100 < 100 is false      -> stay top-down
#  L2 TD  94: (1600-194)/15 = 93.73 < 94               -> bottom-up
#  L3..L7 BU: 1600/25 = 64 > next is false (L7: 64 > 64 false)
#  L8 BU   1: 64 > 1                                   -> top-down for good
#  L9 TD   0 (the last layer's edge leads back to the pivot)
FW_A = [(0, 1, 0, 100), (0, 2, 0, 94), (0, 3, 1, 300), (0, 4, 1, 500), (0, 5, 1, 400), (0, 6, 1, 140),
        (0, 7, 1, 64), (0, 8, 1, 1), (0, 9, 0, 0)]
# BW from the pivot over in-lists: L8 (1), L7 (64), L6 (140): (1600-205)/15 = 93 < 140 -> BU;
# the ascending bottom-up sweep then claims L5 and, through them, L4 .. L1 in one level
# (their out-neighbours were claimed earlier in the same sweep): 400+500+300+94+100 = 1394;
# an empty BU level follows (64 > 0: back to top-down, queue empty).
BW_A = [(1, 1, 0, 1), (1, 2, 0, 64), (1, 3, 0, 140), (1, 4, 1, 1394), (1, 5, 1, 0)]
# n = 1616 (16 isolated vertices, trimmed: num_valid stays 1600):
#  L2: (1616-194)/15 = 94.8 < 94 is false -> top-down (with the valid count 1600 it would switch)
#  L3 TD 300: (1616-494)/15 = 74.8 < 300  -> bottom-up
#  L7 BU 64: 1616/25 = 64 (integer) > 64 is false -> stay (64.64 > 64 would switch back)
FW_B = [(0, 1, 0, 100), (0, 2, 0, 94), (0, 3, 0, 300), (0, 4, 1, 500), (0, 5, 1, 400), (0, 6, 1, 140),
        (0, 7, 1, 64), (0, 8, 1, 1), (0, 9, 0, 0)]
BW_B = BW_A


@pytest.mark.parametrize("isolated,fw,bw", [(0, FW_A, BW_A), (16, FW_B, BW_B)])
def test_oracle_switch_known_answers(msc, orc, isolated, fw, bw):
    g, piv = layered(msc, SIZES, isolated)
    assert g.n == 1600 + isolated
    orc.set_threads(1)  # one thread: the bottom-up sweep runs in ascending id order
    try:
        _, lv, summ = orc.scc(g, piv, 1.0, vert_cutoff=0, trace_cap=100)
    finally:
        orc.set_threads(8)
    assert [x for x in lv if x[0] == 0] == fw
    assert [x for x in lv if x[0] == 1] == bw
    assert summ["num_trim"] == isolated and summ["num_fwbw"] == 1600


def _until_first_bottom_up(trace):
    for k, x in enumerate(trace):
        if x[2] == 1:
            return k
    return None


def _cases(msc):
    ga, pa = layered(msc, SIZES, 0)
    gb, pb = layered(msc, SIZES, 16)
    yield "layered", ga, pa, 1.0
    yield "layered+isolated", gb, pb, 1.0
    for s in (12, 16, 18):
        g = msc.gen_rmat(s, 16, seed=s)
        yield f"rmat{s}", g, g.max_deg_vert, g.avg_degree
    for side in (256, 1024):
        g = msc.gen_grid(side, side, seed=3)
        yield f"grid{side}", g, g.max_deg_vert, g.avg_degree
    g = msc.gen_planted(20000, zipf_s=1.5, max_size=5000, seed=4)[0]
    yield "planted", g, g.max_deg_vert, g.avg_degree


@pytest.mark.gpu
@pytest.mark.parametrize("concurrent", [0, 1, 2])
def test_plain_schedule_matches_oracle(msc, orc, gpu_ctx, concurrent):
    defaults = gpu_ctx.tuning()
    try:
        gpu_ctx.set_tuning(concurrent_fwbw=concurrent)
        for name, g, piv, avg in _cases(msc):
            print(f"case: plain schedule {name} concurrent_fwbw={concurrent} n={g.n} m={g.m}", flush=True)
            got, st = msc.run_scc(g, piv, avg, labels=msc.LABELS_RAW, ctx=gpu_ctx, record=True, stats=True)
            recs = gpu_ctx.levels()
            want, lv, _ = orc.scc(g, piv, avg, vert_cutoff=0, trace_cap=1 << 14)
            assert np.array_equal(got, want), name
            # beside the forward search, the backward one searches the whole valid set
            # (SCC = FW set & BW set; with 2 until the forward search has finished), so
            # only FW follows the reference's schedule there
            for phase in ((0, 1) if concurrent == 0 else (0,)):
                dev = [(r.phase, r.level, r.bottom_up, r.next_size) for r in recs if r.phase == phase]
                ref = [x for x in lv if x[0] == phase]
                k = _until_first_bottom_up(ref)
                if k is None:
                    assert dev == ref, (name, phase)
                else:
                    assert dev[:k] == ref[:k], (name, phase)
                    assert dev[k][:3] == ref[k][:3], (name, phase, "first bottom-up decision")
    finally:
        gpu_ctx.set_tuning(**defaults)
