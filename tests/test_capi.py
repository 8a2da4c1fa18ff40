"""The C ABI (include/msc.h): every declared entry point is exported by libmsc.so, the
struct layouts the Python layer uses match the header, and argument errors come back as
status codes.  No GPU needed (nothing here launches a kernel)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "msc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(msc_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(msc):
    lib = ctypes.CDLL(msc.lib_path())
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing
    assert len(declared_functions()) >= 25


def test_no_oracle_or_reference_linked(msc):
    out = subprocess.run(["ldd", msc.lib_path()], capture_output=True, text=True).stdout
    assert "oracle" not in out and "ref_" not in out
    syms = subprocess.run(["nm", "-D", msc.lib_path()], capture_output=True, text=True).stdout
    assert "orc_" not in syms and "ref_scc" not in syms


def test_struct_sizes_match_header(msc, tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "msc.h"\n'
                   "int main(){printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(msc_csr), sizeof(msc_opts),"
                   " sizeof(msc_stats), sizeof(msc_verify), sizeof(msc_level_rec), sizeof(msc_host_graph));}")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(t) for t in (msc._Csr, msc._Opts, msc._Stats, msc._Verify, msc._LevelRec, msc._HostGraph)]
    assert got == want


def test_opts_defaults_are_the_reference_constants(msc):
    o = msc._Opts()
    msc._lib().msc_opts_default(ctypes.byref(o))
    # ALPHA 15.0, BETA 25, TRIM_LEVEL 1 (multistep/scc_main.cpp:61-65)
    assert (o.alpha, o.beta, o.trim_level, o.policy, o.label_mode) == (15.0, 25, 1, 0, 0)


def test_argument_errors_are_status_codes(msc):
    L = msc._lib()
    o = msc._Opts()
    L.msc_opts_default(ctypes.byref(o))
    assert L.msc_scc_run(None, -1, -1.0, None, ctypes.byref(o), None) == msc.E_INVALID
    g = msc.create_csr(3, [0, 1], [1, 2])
    csr = g.csr()
    lab = np.zeros(3, np.int32)
    assert L.msc_scc_run(ctypes.byref(csr), 7, -1.0, lab.ctypes.data, ctypes.byref(o), None) == msc.E_INVALID
    bad = msc._Opts()
    L.msc_opts_default(ctypes.byref(bad))
    bad.trim_level = 3  # 1 (scc_trim) and 2 (scc_trim_complete) exist; CC takes only 1
    assert L.msc_scc_run(ctypes.byref(csr), -1, -1.0, lab.ctypes.data, ctypes.byref(bad), None) == msc.E_INVALID
    bad.trim_level = 2
    assert L.msc_cc_run(ctypes.byref(csr), -1, -1.0, lab.ctypes.data, ctypes.byref(bad), None) == msc.E_INVALID
    bad.trim_level, bad.beta = 1, 0
    assert L.msc_scc_run(ctypes.byref(csr), -1, -1.0, lab.ctypes.data, ctypes.byref(bad), None) == msc.E_INVALID
    # malformed offsets are refused before any device work
    broken = msc._Csr(3, 2, g.out_off.ctypes.data, g.out_adj.ctypes.data, g.in_off.ctypes.data,
                      g.in_adj.ctypes.data, 0)
    oo = g.out_off.copy()
    oo[3] = 5
    broken.out_off = oo.ctypes.data
    assert L.msc_scc_run(ctypes.byref(broken), -1, -1.0, lab.ctypes.data, ctypes.byref(o), None) == msc.E_INVALID


def test_noalloc_refuses_what_would_allocate(msc):
    # msc_scc_run_noalloc needs a context made beforehand and device-resident graph and
    # labels; each refusal comes back before any device call
    L = msc._lib()
    o = msc._Opts()
    L.msc_opts_default(ctypes.byref(o))
    g = msc.create_csr(3, [0, 1], [1, 2])
    csr = g.csr()
    lab = np.zeros(3, np.int32)
    assert L.msc_scc_run_noalloc(None, -1, -1.0, None, ctypes.byref(o), None) == msc.E_INVALID
    # no context (the default context would be created, i.e. allocated, on the fly)
    assert L.msc_scc_run_noalloc(ctypes.byref(csr), -1, -1.0, lab.ctypes.data, ctypes.byref(o), None) == msc.E_INVALID
    o.labels_where = msc.DEVICE
    assert L.msc_scc_run_noalloc(ctypes.byref(csr), -1, -1.0, lab.ctypes.data, ctypes.byref(o), None) == msc.E_INVALID


def test_no_device_is_reported_not_hidden(msc):
    if msc.device_count() > 0:
        pytest.skip("a GPU is visible")
    g = msc.create_csr(3, [0, 1], [1, 2])
    with pytest.raises(msc.MscError) as e:
        msc.run_scc(g)
    assert e.value.code == msc.E_DEVICE


def test_host_utilities(msc, tmp_path):
    with pytest.raises(msc.MscError):
        msc.create_csr(3, [0, 5], [1, 2])  # id out of range
    with pytest.raises(msc.MscError):
        msc.read_edge(str(tmp_path / "missing.txt"))
    p = tmp_path / "g.txt"
    p.write_text("4 3\n0 1\n1 0\n2 3\n")
    g = msc.read_edge(str(p))
    assert (g.n, g.m) == (4, 3)
    lab = np.array([0, 0, 2, 3], np.int32)
    msc.write_labels(str(tmp_path / "o.txt"), lab)
    assert (tmp_path / "o.txt").read_text() == "0\n0\n2\n3\n"
    assert list(msc.canonicalize([5, 5, 2, 2, -1, 5])) == [0, 0, 2, 2, -1, 0]
    assert msc.cc_verify([0, 0, 2]) == (2, 2, 0)
