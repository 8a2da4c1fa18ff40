"""ORACLE -- test infrastructure only.

ctypes loader for the CPU restatement of the reference's Multistep SCC / CC
(``multistep_oracle.cpp``, built into ``_build/liboracle.so``) and for the reference's own
multistep sources compiled by ``Makefile`` into ``_ref/`` (present only where
HPCGraphAnalysis/Connectivity was available at build time).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this
module, and only as the checker / CPU baseline.  The product (``connectivity_amd``) never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SCC_SO = os.path.join(HERE, "_ref", "libref_scc.so")
REF_CC_SO = os.path.join(HERE, "_ref", "libref_cc.so")
REF_SCC_BIN = os.path.join(HERE, "_ref", "scc")
REF_CC_BIN = os.path.join(HERE, "_ref", "cc")

PHASE_FW, PHASE_BW, PHASE_CC = 0, 1, 2


class _Level(C.Structure):
    _fields_ = [("phase", C.c_int32), ("level", C.c_int32), ("bottom_up", C.c_int32), ("next_size", C.c_int32)]


class _Trace(C.Structure):
    _fields_ = [("levels", C.POINTER(_Level)), ("cap", C.c_int32), ("count", C.c_int32),
                ("num_trim", C.c_int32), ("num_fwbw", C.c_int32), ("num_color", C.c_int32),
                ("num_serial", C.c_int32), ("color_rounds", C.c_int32), ("color_iters", C.c_int32)]


_lib = None
_ref_scc = None
_ref_cc = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle oracle`")
        L = C.CDLL(ORACLE_SO)
        P = C.c_void_p
        L.orc_pivot_scc.argtypes = [C.c_int32, P, P, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.orc_pivot_cc.argtypes = [C.c_int32, P, P, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.orc_scc_run.argtypes = [C.c_int32, P, P, P, P, C.c_int32, C.c_double, C.c_int32, C.c_double,
                                  C.c_int32, P, C.POINTER(_Trace)]
        L.orc_scc_run_trim.argtypes = [C.c_int32, P, P, P, P, C.c_int32, C.c_double, C.c_int32, C.c_double,
                                       C.c_int32, C.c_int32, P, C.POINTER(_Trace)]
        L.orc_trim_complete.argtypes = [C.c_int32, P, P, P, P, P]
        L.orc_cc_run.argtypes = [C.c_int32, P, P, P, P, C.c_int32, C.c_double, C.c_double, C.c_int32, P,
                                 C.POINTER(_Trace)]
        L.orc_canonicalize.argtypes = [C.c_int32, P]
        L.orc_verify_counts.argtypes = [C.c_int32, P, P]
        L.orc_tarjan.argtypes = [C.c_int32, P, P, P]
        L.orc_wcc_unionfind.argtypes = [C.c_int32, P, P, P]
        L.orc_scc_time.argtypes = [C.c_int32, P, P, P, P, C.c_int32, C.c_double, C.c_int32, P]
        L.orc_scc_time.restype = C.c_double
        _lib = L
    return _lib


def set_threads(k: int) -> None:
    """OpenMP threads of the calling thread's next parallel regions (omp_set_num_threads)."""
    C.CDLL("libgomp.so.1").omp_set_num_threads(int(k))


def have_ref() -> bool:
    return os.path.exists(REF_SCC_SO) and os.path.exists(REF_CC_SO)


def ref_scc():
    global _ref_scc
    if _ref_scc is None:
        L = C.CDLL(REF_SCC_SO)
        P = C.c_void_p
        L.ref_scc_run.argtypes = [C.c_int, C.c_uint, P, P, P, P, C.c_int, C.c_double, C.c_int, P]
        L.ref_scc_time.argtypes = [C.c_int, C.c_uint, P, P, P, P, C.c_int, C.c_double, C.c_int, P]
        L.ref_scc_time.restype = C.c_double
        L.ref_create_csr.argtypes = [C.c_int, C.c_uint, P, P, P, P, P, P, C.POINTER(C.c_int),
                                     C.POINTER(C.c_double)]
        L.ref_trim_complete.argtypes = [C.c_int, C.c_uint, P, P, P, P, P]
        _ref_scc = L
    return _ref_scc


def ref_cc():
    global _ref_cc
    if _ref_cc is None:
        L = C.CDLL(REF_CC_SO)
        P = C.c_void_p
        L.ref_cc_run.argtypes = [C.c_int, C.c_uint, P, P, P, P]
        L.ref_cc_time.argtypes = [C.c_int, C.c_uint, P, P, P]
        L.ref_cc_time.restype = C.c_double
        _ref_cc = L
    return _ref_cc


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


def _arrays(g):
    return (np.ascontiguousarray(g.out_off, np.uint32), np.ascontiguousarray(g.out_adj, np.int32),
            np.ascontiguousarray(g.in_off, np.uint32), np.ascontiguousarray(g.in_adj, np.int32))


def pivot_scc(g):
    oo, _, io, _ = _arrays(g)
    piv, avg = C.c_int32(), C.c_double()
    lib().orc_pivot_scc(g.n, _p(oo), _p(io), C.byref(piv), C.byref(avg))
    return piv.value, avg.value


def pivot_cc(g):
    oo, _, io, _ = _arrays(g)
    piv, avg = C.c_int32(), C.c_double()
    lib().orc_pivot_cc(g.n, _p(oo), _p(io), C.byref(piv), C.byref(avg))
    return piv.value, avg.value


def trim_complete(g):
    """The complete trim (TRIM_LEVEL 2) alone: v for trimmed vertices, -1 for survivors."""
    oo, oa, io, ia = _arrays(g)
    labels = np.empty(g.n, dtype=np.int32)
    if lib().orc_trim_complete(g.n, _p(oo), _p(oa), _p(io), _p(ia), _p(labels)) < 0:
        raise ValueError("orc_trim_complete failed")
    return labels


def scc(g, pivot=None, avg=None, vert_cutoff=10000, alpha=15.0, beta=25, trace_cap=0, trim_level=1):
    """Raw labels of the restated run_scc; with trace_cap > 0 also (levels, summary)."""
    oo, oa, io, ia = _arrays(g)
    if pivot is None or avg is None:
        p, a = pivot_scc(g)
        pivot = p if pivot is None else pivot
        avg = a if avg is None else avg
    labels = np.empty(g.n, dtype=np.int32)
    tr = _Trace()
    buf = (_Level * trace_cap)() if trace_cap else None
    tr.levels = buf if trace_cap else None
    tr.cap = trace_cap
    rc = lib().orc_scc_run_trim(g.n, _p(oo), _p(oa), _p(io), _p(ia), int(pivot), float(avg), int(vert_cutoff),
                                float(alpha), int(beta), int(trim_level), _p(labels), C.byref(tr))
    if rc != 0:
        raise ValueError(f"orc_scc_run failed ({rc})")
    if trace_cap:
        lv = [(buf[i].phase, buf[i].level, buf[i].bottom_up, buf[i].next_size) for i in range(min(tr.count, trace_cap))]
        summ = {k: getattr(tr, k) for k in ("num_trim", "num_fwbw", "num_color", "num_serial", "color_rounds",
                                             "color_iters")}
        return labels, lv, summ
    return labels


def cc(g, pivot=None, avg=None, alpha=15.0, beta=25, trace_cap=0):
    """Raw labels of the restated run_cc (weak components of the digraph)."""
    oo, oa, io, ia = _arrays(g)
    if pivot is None or avg is None:
        p, a = pivot_cc(g)
        pivot = p if pivot is None else pivot
        avg = a if avg is None else avg
    labels = np.empty(g.n, dtype=np.int32)
    tr = _Trace()
    buf = (_Level * trace_cap)() if trace_cap else None
    tr.levels = buf if trace_cap else None
    tr.cap = trace_cap
    rc = lib().orc_cc_run(g.n, _p(oo), _p(oa), _p(io), _p(ia), int(pivot), float(avg), float(alpha), int(beta),
                          _p(labels), C.byref(tr))
    if rc != 0:
        raise ValueError(f"orc_cc_run failed ({rc})")
    if trace_cap:
        lv = [(buf[i].phase, buf[i].level, buf[i].bottom_up, buf[i].next_size) for i in range(min(tr.count, trace_cap))]
        summ = {k: getattr(tr, k) for k in ("num_trim", "num_fwbw", "num_color", "color_rounds", "color_iters")}
        return labels, lv, summ
    return labels


def canonical(labels):
    lab = np.array(labels, dtype=np.int32, copy=True)
    lib().orc_canonicalize(lab.shape[0], _p(lab))
    return lab


def verify_counts(labels):
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    out = np.zeros(4, dtype=np.int32)
    lib().orc_verify_counts(lab.shape[0], _p(lab), _p(out))
    return tuple(int(x) for x in out)


def tarjan(g):
    oo, oa, _, _ = _arrays(g)
    labels = np.empty(g.n, dtype=np.int32)
    lib().orc_tarjan(g.n, _p(oo), _p(oa), _p(labels))
    return labels


def wcc(g):
    oo, oa, _, _ = _arrays(g)
    labels = np.empty(g.n, dtype=np.int32)
    lib().orc_wcc_unionfind(g.n, _p(oo), _p(oa), _p(labels))
    return labels


def scc_time(g, pivot, avg, vert_cutoff=10000):
    oo, oa, io, ia = _arrays(g)
    return lib().orc_scc_time(g.n, _p(oo), _p(oa), _p(io), _p(ia), int(pivot), float(avg), int(vert_cutoff), None)


# ---- the reference's own code (oracle/_ref) -------------------------------------------------
def ref_run_scc(g, pivot, avg, vert_cutoff=10000):
    oo, oa, io, ia = _arrays(g)
    labels = np.empty(g.n, dtype=np.int32)
    ref_scc().ref_scc_run(g.n, g.m, _p(oo), _p(oa), _p(io), _p(ia), int(pivot), float(avg), int(vert_cutoff),
                          _p(labels))
    return labels


def ref_time_scc(g, pivot, avg, vert_cutoff=10000):
    oo, oa, io, ia = _arrays(g)
    return ref_scc().ref_scc_time(g.n, g.m, _p(oo), _p(oa), _p(io), _p(ia), int(pivot), float(avg),
                                  int(vert_cutoff), None)


def ref_trim_complete(g):
    """The reference's scc_trim_complete on a fresh valid set (v trimmed, -1 survivor)."""
    oo, oa, io, ia = _arrays(g)
    labels = np.empty(g.n, dtype=np.int32)
    ref_scc().ref_trim_complete(g.n, g.m, _p(oo), _p(oa), _p(io), _p(ia), _p(labels))
    return labels


def ref_create_csr(n, srcs, dsts):
    s = np.ascontiguousarray(srcs, dtype=np.int32)
    d = np.ascontiguousarray(dsts, dtype=np.int32)
    m = s.shape[0]
    oo = np.empty(n + 1, np.uint32)
    io = np.empty(n + 1, np.uint32)
    oa = np.empty(max(m, 1), np.int32)
    ia = np.empty(max(m, 1), np.int32)
    piv, avg = C.c_int(-1), C.c_double()
    ref_scc().ref_create_csr(n, m, _p(s), _p(d), _p(oo), _p(oa), _p(io), _p(ia), C.byref(piv), C.byref(avg))
    return oo, oa[:m], io, ia[:m], piv.value, avg.value


def ref_run_cc(n, srcs, dsts):
    s = np.ascontiguousarray(srcs, dtype=np.int32)
    d = np.ascontiguousarray(dsts, dtype=np.int32)
    labels = np.empty(n, dtype=np.int32)
    piv = C.c_int(-1)
    ref_cc().ref_cc_run(n, s.shape[0], _p(s), _p(d), _p(labels), C.byref(piv))
    return labels


def ref_time_cc(n, srcs, dsts):
    s = np.ascontiguousarray(srcs, dtype=np.int32)
    d = np.ascontiguousarray(dsts, dtype=np.int32)
    return ref_cc().ref_cc_time(n, s.shape[0], _p(s), _p(d), None)
