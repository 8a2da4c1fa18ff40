"""Bounds of every index a pull level forms (pull_group, connectivity_amd/csrc/msc_kernels.hip),
checked on the CPU over the benchmark inputs: the 4096 x 4096 grid digraph and the batch of
8192 graphs, SCC (in-lists) and mirrored WCC (out- and in-lists), with every vertex valid
(every bitmap word full, the most candidates a group can have), the degree-trim survivors,
and a random half.  tests/pull_bounds.cpp restates the kernel's arithmetic: the XCD-dealt
group positions, the candidates' list bounds, the binary lifting over their prefix, list
indices, neighbour ids, LDS slots and the stage bound.  No GPU: these are the indices a
device fault in this code would need, so a clean run rules out an out-of-bounds access in
pull_group for these inputs."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = tmp_path_factory.mktemp("pull") / "libpull_bounds.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", str(out),
                    os.path.join(HERE, "pull_bounds.cpp")], check=True)
    lib = C.CDLL(str(out))
    P = C.c_void_p
    lib.pull_bounds_check.argtypes = [C.c_int32, P, P, P, P, P, P, P, C.c_int64, C.POINTER(C.c_int64)]
    lib.pull_bounds_check.restype = C.c_longlong
    return lib


def _bits(mask):
    n = mask.shape[0]
    nw = (n + 31) // 32
    padded = np.zeros(nw * 32, dtype=np.uint8)
    padded[:n] = mask
    return np.packbits(padded.reshape(-1, 8)[:, ::-1], axis=1).reshape(-1).view("<u4").copy(), nw


def _run(lib, g, mirrored, mask):
    words, nw = _bits(mask)
    p = lambda a: a.ctypes.data if a is not None else None  # noqa: E731
    if mirrored:  # the mirrored pull walks out-lists then in-lists
        args = (g.out_off, g.out_adj, g.in_off, g.in_adj, g.out_off, g.in_off)
    else:
        args = (g.in_off, g.in_adj, None, None, g.out_off, None)
    walked = C.c_int64()
    bad = lib.pull_bounds_check(g.n, *(p(a) for a in args), p(words), nw, C.byref(walked))
    return bad, walked.value


def _masks(g, seed):
    n = g.n
    deg_out, deg_in = np.diff(g.out_off.astype(np.int64)), np.diff(g.in_off.astype(np.int64))
    rng = np.random.default_rng(seed)
    return {"all": np.ones(n, np.uint8), "trimmed": ((deg_out > 0) & (deg_in > 0)).astype(np.uint8),
            "half": (rng.random(n) < 0.5).astype(np.uint8)}


def test_bit_packing_matches_the_device_layout():
    # bit v of word v >> 5 is (v & 31), little-endian words
    mask = np.zeros(70, np.uint8)
    mask[[0, 5, 31, 32, 69]] = 1
    words, nw = _bits(mask)
    assert nw == 3 and list(words) == [(1 << 0) | (1 << 5) | (1 << 31), 1, 1 << 5]


def test_checker_reports_a_planted_bad_id(msc, checker):
    # an adjacency entry outside [0, n) is what a wrong list index would read; the checker
    # must report it (the planted fault is in the input, on the CPU)
    g = msc.gen_rmat(10, 8, seed=3)
    ok, _ = _run(checker, g, False, np.ones(g.n, np.uint8))
    assert ok == 0
    g.in_adj = g.in_adj.copy()
    g.in_adj[g.m // 2] = g.n + 7
    bad, _ = _run(checker, g, False, np.ones(g.n, np.uint8))
    assert bad >= 1


def test_pull_indices_grid4096(msc, checker):
    g = msc.gen_grid(4096, 4096, 0.5, 0.1, 0.005, seed=1)
    for name, mask in _masks(g, 1).items():
        for mirrored in (False, True):
            bad, walked = _run(checker, g, mirrored, mask)
            assert bad == 0, (name, mirrored)
            assert walked > 0


@pytest.mark.slow
def test_pull_indices_batch8192(msc, checker):
    g, base, _ = msc.gen_batch(8192, seed=1)
    assert g.n == int(base[-1])
    for name, mask in _masks(g, 2).items():
        if name == "half":
            continue  # the grid covers a random valid set; here every vertex, then the survivors
        for mirrored in (False, True):
            bad, walked = _run(checker, g, mirrored, mask)
            assert bad == 0, (name, mirrored)
            if name == "all":
                assert walked == (2 * g.m if mirrored else g.m)
