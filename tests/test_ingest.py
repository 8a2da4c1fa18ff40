"""Device-side graph ingest: read_edge + create_csr (multistep/scc_main.cpp:94-191) on the GPU.

The device CSR must equal the host build with the reference's semantics (msc.read_edge,
whose create_csr restatement test_oracle pins against the reference's own create_csr): the
same offsets, the same adjacency in input order (duplicates and self-loops kept), the same
pivot and average degree.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def text_of(n, src, dst, newline="\n", trailing=True):
    body = newline.join(f"{u} {v}" for u, v in zip(src.tolist(), dst.tolist()))
    return (f"{n} {len(src)}{newline}" + body + (newline if trailing else "")).encode()


def test_format_round_trip(msc, tmp_path):
    g = msc.gen_rmat(11, 8, seed=2)
    t = msc.format_edge_list(g)
    assert t.startswith(f"{g.n} {g.m}\n".encode()) and t.endswith(b"\n")
    p = tmp_path / "g.txt"
    p.write_bytes(t)
    h = msc.read_edge(str(p))
    for f in ("out_off", "out_adj", "in_off", "in_adj"):
        assert np.array_equal(getattr(h, f), getattr(g, f))


def same_csr(a, b):
    return all(np.array_equal(getattr(a, f), getattr(b, f)) for f in ("out_off", "out_adj", "in_off", "in_adj"))


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["fixture", "random", "rmat", "crlf", "no_trailing_newline", "extra_lines"])
def test_ingest_matches_host_build(msc, gpu_ctx, tmp_path, case):
    print(f"case: ingest {case}", flush=True)
    rng = np.random.default_rng(5)
    if case == "fixture":
        path = os.path.join(GOLDEN, "test.graph")
        text = open(path, "rb").read()
    else:
        if case == "rmat":
            g0 = msc.gen_rmat(16, 16, seed=3)
            src, dst = g0.edges()
            n = g0.n
        else:
            n = 5000
            m = 40000
            src, dst = rng.integers(0, n, m), rng.integers(0, n, m)
            src[:50] = dst[:50]  # self-loops
            src[50:100], dst[50:100] = src[100:150], dst[100:150]  # duplicates
        order = rng.permutation(len(src))  # input order is not CSR order
        src, dst = src[order], dst[order]
        if case == "crlf":
            text = text_of(n, src, dst, newline="\r\n")
        elif case == "no_trailing_newline":
            text = text_of(n, src, dst, trailing=False)
        elif case == "extra_lines":
            text = text_of(n, src, dst) + b"1 2\n3 4\n"
        else:
            text = text_of(n, src, dst)
        path = str(tmp_path / "g.txt")
        open(path, "wb").write(text)
    host = msc.read_edge(path)
    for dg in (msc.ingest_text(gpu_ctx, text), msc.read_edge_device(gpu_ctx, path)):
        got = dg.to_host()
        assert same_csr(got, host), case
        assert dg.max_deg_vert == host.max_deg_vert and dg.avg_degree == host.avg_degree
        # the device graph feeds the hot path directly
        assert np.array_equal(msc.run_scc(dg, dg.max_deg_vert, dg.avg_degree, ctx=gpu_ctx),
                              msc.run_scc(host, host.max_deg_vert, host.avg_degree, ctx=gpu_ctx))
        dg.free()


@pytest.mark.gpu
def test_ingest_rejects_bad_text(msc, gpu_ctx):
    for bad in (b"3 2\n0 1\n2 3\n",      # id outside [0, n)
                b"3 3\n0 1\n1 2\n",      # fewer lines than m
                b"3 2\n0 1\n12\n",       # no space between the ids
                b"3 2\n-1 1\n1 2\n",     # negative id
                b"x"):                   # no header
        with pytest.raises(msc.MscError):
            msc.ingest_text(gpu_ctx, bad)
    # an empty edge set is a graph of isolated vertices
    dg = msc.ingest_text(gpu_ctx, b"4 0\n")
    g = dg.to_host()
    assert g.n == 4 and g.m == 0 and np.all(g.out_off == 0) and np.all(g.in_off == 0)
    dg.free()
