"""Golden fixtures written by tests/golden/make_golden.py from the reference's own code.

CPU tests: the oracle and the product's host-side pieces (CSR build, generators, verify
lines) reproduce the fixtures.  GPU tests (marked gpu): libmsc on the device reproduces
them through the C ABI.
"""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(__file__), "golden")
NPZ = np.load(os.path.join(HERE, "golden.npz"), allow_pickle=False)
META = json.load(open(os.path.join(HERE, "golden.json")))
SMALL = sorted({k.rsplit("_", 1)[0] for k in NPZ.files if k.endswith("_src")})
GENERATED = {
    "rmat14_seed1": lambda msc: msc.gen_rmat(14, 16, seed=1),
    "rmat16_seed1": lambda msc: msc.gen_rmat(16, 16, seed=1),
    "grid256_seed1": lambda msc: msc.gen_grid(256, 256, seed=1),
    "planted20000_seed4": lambda msc: msc.gen_planted(20000, max_size=5000, seed=4)[0],
}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def graph_hash(g):
    h = hashlib.sha256()
    for a in (g.out_off, g.out_adj, g.in_off, g.in_adj):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def small(msc, name):
    n = int(NPZ[f"{name}_n"][0])
    return msc.create_csr(n, NPZ[f"{name}_src"], NPZ[f"{name}_dst"])


@pytest.mark.parametrize("name", SMALL)
def test_oracle_small(msc, orc, name):
    g = small(msc, name)
    piv = int(NPZ[f"{name}_pivot"][0])
    if piv >= 0:
        assert (g.max_deg_vert, g.avg_degree) == (piv, float(NPZ[f"{name}_avg"][0]))
        assert np.array_equal(orc.scc(g, vert_cutoff=0), NPZ[f"{name}_scc_cut0"])
        assert np.array_equal(orc.scc(g, vert_cutoff=10000), NPZ[f"{name}_scc_cut10000"])
    if f"{name}_cc" in NPZ.files:
        assert np.array_equal(orc.cc(g), NPZ[f"{name}_cc"])


@pytest.mark.parametrize("name", sorted(GENERATED))
def test_generators_and_oracle_hashed(msc, orc, name):
    ent = META["graphs"][name]
    g = GENERATED[name](msc)
    assert (g.n, g.m, g.max_deg_vert, g.avg_degree) == (ent["n"], ent["m"], ent["pivot"], ent["avg_degree"])
    assert graph_hash(g) == ent["csr_sha256"], "generator output changed"
    raw = orc.scc(g, vert_cutoff=0)
    assert sha(raw) == ent["scc_raw_cut0_sha256"]
    assert list(orc.verify_counts(raw)) == ent["scc_counts"]
    if "cc_raw_sha256" in ent:
        assert sha(orc.cc(g)) == ent["cc_raw_sha256"]


def test_verify_lines_match_reference_executables(msc):
    g = msc.read_edge(os.path.join(HERE, "test.graph"))
    tg = META["test_graph"]
    assert msc.verify_line(tg["scc"]["labels"], "scc") == tg["scc"]["summary"]
    assert msc.verify_line(tg["cc"]["labels"], "cc") == tg["cc"]["summary"]
    assert msc.scc_verify(tg["scc"]["labels"]) == (2, 1, 2, 0)
    assert g.n == 3


def test_generated_graph_properties(msc):
    """No self-loops, no duplicates, sorted lists, CSC is the transpose of CSR."""
    for g in (msc.gen_rmat(12, 16, seed=7), msc.gen_grid(50, 70, seed=7), msc.gen_planted(3000, seed=7)[0]):
        src, dst = g.edges()
        assert not np.any(src == dst)
        key = src.astype(np.int64) * g.n + dst
        assert np.all(np.diff(key) > 0)  # strictly increasing: sorted rows, unique
        tsrc = np.repeat(np.arange(g.n), np.diff(g.in_off.astype(np.int64)))
        tkey = g.in_adj.astype(np.int64) * g.n + tsrc
        assert np.array_equal(np.sort(tkey), key)


def test_generator_thread_independence(msc):
    import subprocess
    import sys

    code = ("import sys, hashlib, numpy as np; sys.path.insert(0,'.'); import connectivity_amd as m;"
            "g=m.gen_rmat(13,16,seed=9); h=hashlib.sha256();"
            "[h.update(np.ascontiguousarray(a).tobytes()) for a in (g.out_off,g.out_adj,g.in_off,g.in_adj)];"
            "print(h.hexdigest())")
    outs = {subprocess.run([sys.executable, "-c", code], env={**os.environ, "OMP_NUM_THREADS": t},
                           capture_output=True, text=True, check=True).stdout for t in ("1", "3", "8")}
    assert len(outs) == 1


# ---- GPU ----------------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("name", SMALL)
def test_gpu_small(msc, gpu_ctx, name):
    print(f"case: golden {name}", flush=True)
    g = small(msc, name)
    piv = int(NPZ[f"{name}_pivot"][0])
    if piv >= 0:
        got = msc.run_scc(g, piv, float(NPZ[f"{name}_avg"][0]), labels=msc.LABELS_RAW, ctx=gpu_ctx)
        assert np.array_equal(got, NPZ[f"{name}_scc_cut0"])
    if f"{name}_cc" in NPZ.files:
        assert np.array_equal(msc.run_cc(g, labels=msc.LABELS_RAW, ctx=gpu_ctx), NPZ[f"{name}_cc"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GENERATED))
def test_gpu_hashed(msc, gpu_ctx, name):
    print(f"case: golden {name}", flush=True)
    ent = META["graphs"][name]
    g = GENERATED[name](msc)
    raw = msc.run_scc(g, g.max_deg_vert, g.avg_degree, labels=msc.LABELS_RAW, ctx=gpu_ctx)
    assert sha(raw) == ent["scc_raw_cut0_sha256"]
    assert sha(msc.run_scc(g, ctx=gpu_ctx)) == ent["scc_canonical_sha256"]
    if "cc_raw_sha256" in ent:
        assert sha(msc.run_cc(g, labels=msc.LABELS_RAW, ctx=gpu_ctx)) == ent["cc_raw_sha256"]
        assert sha(msc.run_cc(g, ctx=gpu_ctx)) == ent["cc_canonical_sha256"]
