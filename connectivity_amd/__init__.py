"""Python face of libmsc (MI355X-native Multistep SCC / WCC).

Thin ctypes layer over the C ABI in ``include/msc.h``; the work happens in
``libmsc.so`` (gfx950 HIP kernels + C++ host runtime).  Names mirror the reference's
functions (HPCGraphAnalysis/Connectivity ``multistep/``):

* :func:`read_edge` / :func:`create_csr`  -- read_edge + create_csr
  (multistep/scc_main.cpp:94-191), returning a :class:`Graph`;
* :func:`run_scc`  -- run_scc (multistep/scc_main.cpp:210);
* :func:`run_cc`   -- run_cc (multistep/cc_main.cpp:184) on the digraph's mirror;
* :func:`scc_verify` / :func:`cc_verify` -- the reference's summary counts
  (multistep/scc_verify.cpp:47-77, cc_verify.cpp:47-74).

There is no CPU fallback: every entry point raises :class:`MscError` when the library
is missing or no gfx950 device is present.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "MscError", "Graph", "DeviceGraph", "DeviceLabels", "Context", "Stats", "LevelRecord",
    "read_edge", "create_csr", "gen_rmat", "gen_grid", "gen_planted",
    "run_scc", "run_scc_noalloc", "run_cc", "scc_pivot", "cc_pivot", "batch_pivots", "scc_verify", "cc_verify", "verify_line",
    "canonicalize", "write_labels", "device_count", "lib_path",
    "POLICY_PLAIN", "POLICY_LEARNED", "LABELS_CANONICAL", "LABELS_RAW", "schedule_array",
    "set_policy", "load_policy", "mlp_forward", "mlp_grad", "mlp_train", "POLICY_WEIGHTS",
    "K_TRIM", "K_LEVEL", "K_COMMIT", "K_COLOR_PROP", "K_COLOR_BW", "K_ROOTS", "K_OTHER",
    "KERNEL_NAMES",
]

_HERE = os.path.dirname(os.path.abspath(__file__))

OK, E_INVALID, E_OOM, E_DEVICE, E_CAPACITY, E_IO = 0, -1, -2, -3, -4, -5
_ERRNAME = {E_INVALID: "MSC_E_INVALID", E_OOM: "MSC_E_OOM", E_DEVICE: "MSC_E_DEVICE",
            E_CAPACITY: "MSC_E_CAPACITY", E_IO: "MSC_E_IO"}
HOST, DEVICE = 0, 1
POLICY_PLAIN, POLICY_LEARNED = 0, 1
LABELS_CANONICAL, LABELS_RAW = 0, 1
K_TRIM, K_LEVEL, K_COMMIT, K_COLOR_PROP, K_COLOR_BW, K_ROOTS, K_OTHER = range(7)
KERNEL_NAMES = ["trim", "level", "commit", "color_prop", "color_bw", "roots", "other"]
MLP_IN, MLP_H, MLP_OUT = 12, 32, 3
MLP_PARAMS = MLP_IN * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H * MLP_OUT + MLP_OUT
MLP_BLOB = MLP_PARAMS + 2 * MLP_IN


class MscError(RuntimeError):
    """A negative MSC_E_* status from libmsc."""

    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: {_ERRNAME.get(code, code)}")
        self.code = code


# ---- C structs (layout of include/msc.h) ----------------------------------------------
class _Csr(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int64), ("out_off", C.c_void_p), ("out_adj", C.c_void_p),
                ("in_off", C.c_void_p), ("in_adj", C.c_void_p), ("where", C.c_int32)]


class _HostGraph(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int64), ("out_off", C.POINTER(C.c_uint32)),
                ("out_adj", C.POINTER(C.c_int32)), ("in_off", C.POINTER(C.c_uint32)),
                ("in_adj", C.POINTER(C.c_int32)), ("max_deg_vert", C.c_int32), ("avg_degree", C.c_double)]


class _Opts(C.Structure):
    _fields_ = [("alpha", C.c_double), ("beta", C.c_int32), ("trim_level", C.c_int32),
                ("policy", C.c_int32), ("label_mode", C.c_int32), ("labels_where", C.c_int32),
                ("device", C.c_int32), ("validate", C.c_int32), ("timing", C.c_int32),
                ("record", C.c_int32), ("ctx", C.c_void_p), ("schedule", C.c_void_p),
                ("schedule_len", C.c_int32), ("fwbw_rounds", C.c_int32), ("replay", C.c_int32)]


class _KStat(C.Structure):
    _fields_ = [("launches", C.c_int64), ("ms", C.c_double), ("bytes", C.c_double)]


class _Stats(C.Structure):
    _fields_ = [("wall_ms", C.c_double), ("pivot", C.c_int32), ("avg_degree", C.c_double),
                ("num_trim", C.c_int32), ("num_fwbw", C.c_int32), ("num_color", C.c_int32),
                ("color_rounds", C.c_int32), ("color_iters", C.c_int32),
                ("fw_levels", C.c_int32), ("bw_levels", C.c_int32), ("bu_levels", C.c_int32),
                ("cbw_levels", C.c_int32), ("host_syncs", C.c_int32), ("bytes", C.c_double),
                ("k", _KStat * 7), ("color_tail", C.c_int32)]


class _Verify(C.Structure):
    _fields_ = [("num_components", C.c_int32), ("nontrivial", C.c_int32), ("max_size", C.c_int32),
                ("unassigned", C.c_int32)]


class _LevelRec(C.Structure):
    _fields_ = [("phase", C.c_int32), ("level", C.c_int32), ("bottom_up", C.c_int32),
                ("next_size", C.c_int32), ("q_size", C.c_uint32), ("q_edges", C.c_uint32),
                ("next_edges", C.c_uint32), ("examined", C.c_uint32), ("inspected", C.c_uint64),
                ("found", C.c_int64), ("n_total", C.c_int32), ("num_valid", C.c_int32),
                ("action", C.c_int32), ("pad", C.c_int32), ("feat", C.c_float * 12),
                ("level_us", C.c_float), ("pad2", C.c_int32)]


class _Tuning(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_int32), ("small_push_edges", C.c_int32),
                ("small_cbw_edges", C.c_int32), ("first_batch", C.c_int32), ("max_batch", C.c_int32),
                ("concurrent_fwbw", C.c_int32), ("search_batch", C.c_int32), ("local_rounds", C.c_int32),
                ("local_us", C.c_int32), ("pull_div", C.c_int32),
                ("tail_vertices", C.c_int32), ("balance", C.c_int32), ("group_words", C.c_int32)]


TUNING_FIELDS = ("blocks_per_cu", "small_push_edges", "small_cbw_edges", "first_batch", "max_batch",
                 "concurrent_fwbw", "search_batch", "local_rounds", "local_us", "pull_div",
                 "tail_vertices", "balance", "group_words")


class _TrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("weight_decay", C.c_float), ("seed", C.c_uint64)]


_lib_handle = None


def lib_path() -> str:
    return os.path.join(_HERE, "libmsc.so")


def _lib():
    global _lib_handle
    if _lib_handle is not None:
        return _lib_handle
    path = lib_path()
    if not os.path.exists(path):
        raise MscError(E_DEVICE, f"{path} is not built (run __graft_entry__.build())")
    L = C.CDLL(path)
    P = C.c_void_p
    sig = {
        "msc_opts_default": (None, [C.POINTER(_Opts)]),
        "msc_ctx_create": (C.c_int, [C.c_int32, P, C.POINTER(P)]),
        "msc_ctx_destroy": (C.c_int, [P]),
        "msc_ctx_levels": (C.c_int, [P, C.POINTER(_LevelRec), C.c_int32, C.POINTER(C.c_int32)]),
        "msc_ctx_get_tuning": (C.c_int, [P, C.POINTER(_Tuning)]),
        "msc_ctx_set_tuning": (C.c_int, [P, C.POINTER(_Tuning)]),
        "msc_scc_run": (C.c_int, [C.POINTER(_Csr), C.c_int32, C.c_double, P, C.POINTER(_Opts),
                                  C.POINTER(_Stats)]),
        "msc_cc_run": (C.c_int, [C.POINTER(_Csr), C.c_int32, C.c_double, P, C.POINTER(_Opts),
                                 C.POINTER(_Stats)]),
        "msc_scc_run_noalloc": (C.c_int, [C.POINTER(_Csr), C.c_int32, C.c_double, P, C.POINTER(_Opts),
                                          C.POINTER(_Stats)]),
        "msc_scc_run_batch": (C.c_int, [C.POINTER(_Csr), C.c_int32, P, P, P, C.POINTER(_Opts),
                                        C.POINTER(_Stats)]),
        "msc_cc_run_batch": (C.c_int, [C.POINTER(_Csr), C.c_int32, P, P, P, C.POINTER(_Opts),
                                       C.POINTER(_Stats)]),
        "msc_batch_pack": (C.c_int, [C.c_int32, P, C.POINTER(C.POINTER(_HostGraph)), P]),
        "msc_gen_batch": (C.c_int, [C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_uint64,
                                    C.POINTER(C.POINTER(_HostGraph)), P, P]),
        "msc_scc_pivot": (C.c_int, [C.POINTER(_Csr), C.POINTER(_Opts), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_double)]),
        "msc_cc_pivot": (C.c_int, [C.POINTER(_Csr), C.POINTER(_Opts), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_double)]),
        "msc_scc_verify": (C.c_int, [C.c_int32, P, C.POINTER(_Verify)]),
        "msc_cc_verify": (C.c_int, [C.c_int32, P, C.POINTER(_Verify)]),
        "msc_verify_format": (C.c_int, [C.POINTER(_Verify), C.c_int32, C.c_char_p, C.c_int32]),
        "msc_canonicalize": (C.c_int, [C.c_int32, P]),
        "msc_graph_read_edge_list": (C.c_int, [C.c_char_p, C.POINTER(C.POINTER(_HostGraph))]),
        "msc_graph_from_edges": (C.c_int, [C.c_int32, C.c_int64, P, P, C.c_int32,
                                           C.POINTER(C.POINTER(_HostGraph))]),
        "msc_write_labels": (C.c_int, [C.c_char_p, C.c_int32, P]),
        "msc_graph_free": (None, [C.POINTER(_HostGraph)]),
        "msc_gen_rmat": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                   C.c_uint64, C.c_int32, C.POINTER(C.POINTER(_HostGraph))]),
        "msc_gen_grid": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                   C.c_uint64, C.POINTER(C.POINTER(_HostGraph))]),
        "msc_gen_planted": (C.c_int, [C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_int64,
                                      C.c_uint64, C.POINTER(C.POINTER(_HostGraph)), P]),
        "msc_csr_to_device": (C.c_int, [P, C.POINTER(_Csr), C.POINTER(_Csr)]),
        "msc_graph_ingest_text": (C.c_int, [P, C.c_char_p, C.c_int64, C.POINTER(_Csr), C.POINTER(C.c_int32),
                                            C.POINTER(C.c_double)]),
        "msc_graph_read_edge_list_device": (C.c_int, [P, C.c_char_p, C.POINTER(_Csr), C.POINTER(C.c_int32),
                                                      C.POINTER(C.c_double)]),
        "msc_format_edge_list": (C.c_int, [C.POINTER(_HostGraph), P, C.c_int64, C.POINTER(C.c_int64)]),
        "msc_csr_device_free": (C.c_int, [P, C.POINTER(_Csr)]),
        "msc_version": (C.c_char_p, []),
        "msc_device_count": (C.c_int, []),
        "msc_device_alloc": (C.c_int, [P, C.c_int64, C.POINTER(P)]),
        "msc_device_free": (C.c_int, [P, P]),
        "msc_copy_to_host": (C.c_int, [P, P, P, C.c_int64]),
        "msc_copy_to_device": (C.c_int, [P, P, P, C.c_int64]),
        "msc_copy_device": (C.c_int, [P, P, P, C.c_int64]),
        "msc_ctx_sync": (C.c_int, [P]),
        "msc_policy_set_weights": (C.c_int, [P, P, C.c_int32]),
        "msc_mlp_forward": (C.c_int, [P, P, P, C.c_int64, P]),
        "msc_mlp_grad": (C.c_int, [P, P, P, P, P, C.c_int64, C.POINTER(C.c_float), P]),
        "msc_mlp_train": (C.c_int, [P, P, P, P, P, C.c_int64, C.POINTER(_TrainOpts), P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib_handle = L
    return L


def _check(rc: int, what: str) -> None:
    if rc != OK:
        raise MscError(rc, what)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def device_count() -> int:
    """Visible gfx950 devices (0 when none)."""
    return int(_lib().msc_device_count())


# ---- graphs -----------------------------------------------------------------------------
@dataclass
class Graph:
    """CSR + transposed CSR in host memory (the reference's ``graph`` struct,
    multistep/scc_main.cpp:75-86) plus create_csr's pivot and average degree."""

    n: int
    out_off: np.ndarray  # uint32[n+1]
    out_adj: np.ndarray  # int32[m]
    in_off: np.ndarray   # uint32[n+1]
    in_adj: np.ndarray   # int32[m]
    max_deg_vert: int = -1
    avg_degree: float = 0.0
    _keep: object = field(default=None, repr=False)

    @property
    def m(self) -> int:
        return int(self.out_adj.shape[0])

    def csr(self) -> _Csr:
        return _Csr(self.n, self.m, _ptr(self.out_off), _ptr(self.out_adj), _ptr(self.in_off),
                    _ptr(self.in_adj), HOST)

    def edges(self) -> tuple[np.ndarray, np.ndarray]:
        """(src, dst) arrays in CSR order."""
        deg = np.diff(self.out_off.astype(np.int64))
        return np.repeat(np.arange(self.n, dtype=np.int32), deg), self.out_adj.copy()


class _Owned:
    def __init__(self, ptr):
        self.ptr = ptr

    def __del__(self):
        try:
            _lib().msc_graph_free(self.ptr)
        except Exception:  # interpreter shutdown
            pass


def _wrap(hg_ptr) -> Graph:
    hg = hg_ptr.contents
    n, m = int(hg.n), int(hg.m)
    owner = _Owned(hg_ptr)

    def arr(p, count, dt):
        if count == 0:
            return np.zeros(0, dtype=dt)
        return np.ctypeslib.as_array(p, shape=(count,)).view(dt)

    return Graph(n=n, out_off=arr(hg.out_off, n + 1, np.uint32), out_adj=arr(hg.out_adj, m, np.int32),
                 in_off=arr(hg.in_off, n + 1, np.uint32), in_adj=arr(hg.in_adj, m, np.int32),
                 max_deg_vert=int(hg.max_deg_vert), avg_degree=float(hg.avg_degree), _keep=owner)


BUILD_DROP_SELF_LOOPS, BUILD_DEDUPE = 1, 2


def create_csr(n: int, srcs, dsts, dedupe: bool = False, drop_self_loops: bool = False) -> Graph:
    """create_csr (multistep/scc_main.cpp:127-191).  Without flags, adjacency keeps the input
    order and duplicates, exactly like the reference."""
    s = np.ascontiguousarray(srcs, dtype=np.int32)
    d = np.ascontiguousarray(dsts, dtype=np.int32)
    if s.shape != d.shape:
        raise ValueError("srcs and dsts differ in length")
    flags = (BUILD_DEDUPE if dedupe else 0) | (BUILD_DROP_SELF_LOOPS if drop_self_loops else 0)
    out = C.POINTER(_HostGraph)()
    _check(_lib().msc_graph_from_edges(int(n), int(s.shape[0]), _ptr(s), _ptr(d), flags, C.byref(out)),
           "create_csr")
    return _wrap(out)


def read_edge(path: str) -> Graph:
    """read_edge + create_csr (multistep/scc_main.cpp:94-191): "n m" then m lines "u v"."""
    out = C.POINTER(_HostGraph)()
    _check(_lib().msc_graph_read_edge_list(os.fsencode(path), C.byref(out)), f"read_edge({path})")
    return _wrap(out)


def gen_rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1, permute: bool = True) -> Graph:
    """Directed R-MAT, self-loops and duplicates removed, ids permuted from the seed."""
    out = C.POINTER(_HostGraph)()
    _check(_lib().msc_gen_rmat(scale, edge_factor, a, b, c, seed, 1 if permute else 0, C.byref(out)),
           "gen_rmat")
    return _wrap(out)


def gen_grid(rows: int = 4096, cols: int = 4096, p_one: float = 0.5, p_both: float = 0.1,
             arc_frac: float = 0.005, seed: int = 1) -> Graph:
    """rows x cols grid digraph: each grid edge kept in one random direction with p_one, in
    both with p_both, plus round(arc_frac * n) uniform random arcs."""
    out = C.POINTER(_HostGraph)()
    _check(_lib().msc_gen_grid(rows, cols, p_one, p_both, arc_frac, seed, C.byref(out)), "gen_grid")
    return _wrap(out)


def gen_planted(n: int, zipf_s: float = 1.5, max_size: int = 1 << 20, extra_per_vertex: float = 1.0,
                dag_arcs: int | None = None, seed: int = 1) -> tuple[Graph, np.ndarray]:
    """Planted SCCs (Zipf sizes) joined by random DAG arcs; returns (graph, planted partition)."""
    if dag_arcs is None:
        dag_arcs = 4 * n
    comp = np.empty(n, dtype=np.int32)
    out = C.POINTER(_HostGraph)()
    _check(_lib().msc_gen_planted(n, zipf_s, max_size, extra_per_vertex, dag_arcs, seed, C.byref(out),
                                  _ptr(comp)), "gen_planted")
    return _wrap(out), comp


# ---- contexts and device graphs ------------------------------------------------------------
class Context:
    """A device + HIP stream + reusable workspace (msc_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(_lib().msc_ctx_create(device, None, C.byref(h)), f"msc_ctx_create(device={device})")
        self.handle = h
        self.device = device

    def close(self) -> None:
        if self.handle:
            _lib().msc_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tuning(self) -> dict:
        """Launch parameters of this context (msc_tuning) as a dict."""
        t = _Tuning()
        _check(_lib().msc_ctx_get_tuning(self.handle, C.byref(t)), "msc_ctx_get_tuning")
        return {f: getattr(t, f) for f in TUNING_FIELDS}

    def set_tuning(self, **kw) -> dict:
        """Changes the named launch parameters; the others keep their values."""
        cur = self.tuning()
        bad = set(kw) - set(TUNING_FIELDS)
        if bad:
            raise MscError(E_INVALID, f"unknown tuning fields {sorted(bad)}")
        cur.update(kw)
        t = _Tuning(**{f: int(cur[f]) for f in TUNING_FIELDS})
        _check(_lib().msc_ctx_set_tuning(self.handle, C.byref(t)), f"msc_ctx_set_tuning({cur})")
        return cur

    def levels(self) -> list["LevelRecord"]:
        """Per-level records of the last run made with ``record=True``."""
        count = C.c_int32()
        _check(_lib().msc_ctx_levels(self.handle, None, 0, C.byref(count)), "msc_ctx_levels")
        k = count.value
        if k == 0:
            return []
        buf = (_LevelRec * k)()
        _check(_lib().msc_ctx_levels(self.handle, buf, k, C.byref(count)), "msc_ctx_levels")
        out = []
        for r in buf:
            vals = {f: getattr(r, f) for f, _ in _LevelRec._fields_ if f not in ("pad", "pad2", "feat")}
            out.append(LevelRecord(feat=list(r.feat), **vals))
        return out


class DeviceGraph:
    """A Graph copied into device memory (msc_csr_to_device); runs use it in place.  Also
    built on the device from edge-list text (:func:`ingest_text`, :func:`read_edge_device`)."""

    def __init__(self, ctx: Context, g: Graph | None):
        self.ctx = ctx
        self._csr = _Csr()
        if g is None:
            return
        self.n, self.m = g.n, g.m
        self.max_deg_vert, self.avg_degree = g.max_deg_vert, g.avg_degree
        host = g.csr()
        _check(_lib().msc_csr_to_device(ctx.handle, C.byref(host), C.byref(self._csr)), "msc_csr_to_device")

    def csr(self) -> _Csr:
        return self._csr

    def to_host(self) -> Graph:
        """Copies the four CSR arrays back (for checks)."""
        c = self._csr
        arrs = []
        for ptr, count, dt in ((c.out_off, c.n + 1, np.uint32), (c.out_adj, c.m, np.int32),
                               (c.in_off, c.n + 1, np.uint32), (c.in_adj, c.m, np.int32)):
            a = np.empty(count, dtype=dt)
            if count:
                _check(_lib().msc_copy_to_host(self.ctx.handle, _ptr(a), ptr, a.nbytes), "msc_copy_to_host")
            arrs.append(a)
        return Graph(n=c.n, out_off=arrs[0], out_adj=arrs[1], in_off=arrs[2], in_adj=arrs[3],
                     max_deg_vert=self.max_deg_vert, avg_degree=self.avg_degree)

    def free(self) -> None:
        if self._csr.where == DEVICE and self.ctx.handle:
            _lib().msc_csr_device_free(self.ctx.handle, C.byref(self._csr))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceLabels:
    """n int32 labels kept in device memory (runs write them with labels_where = device)."""

    def __init__(self, ctx: Context, n: int):
        self.ctx, self.n = ctx, n
        self.ptr = C.c_void_p()
        _check(_lib().msc_device_alloc(ctx.handle, 4 * max(n, 1), C.byref(self.ptr)), "msc_device_alloc")

    def to_host(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        _check(_lib().msc_copy_to_host(self.ctx.handle, _ptr(out), self.ptr, 4 * self.n), "msc_copy_to_host")
        return out

    def free(self) -> None:
        if self.ptr and self.ctx.handle:
            _lib().msc_device_free(self.ctx.handle, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class LevelRecord:
    phase: int
    level: int
    bottom_up: int
    next_size: int
    q_size: int
    q_edges: int
    next_edges: int
    examined: int
    inspected: int
    found: int
    n_total: int
    num_valid: int
    action: int
    level_us: float
    feat: list


@dataclass
class Stats:
    wall_ms: float
    pivot: int
    avg_degree: float
    num_trim: int
    num_fwbw: int
    num_color: int
    color_rounds: int
    color_iters: int
    fw_levels: int
    bw_levels: int
    bu_levels: int
    cbw_levels: int
    host_syncs: int
    bytes: float
    kernels: dict
    color_tail: int

    @classmethod
    def _from(cls, s: _Stats) -> "Stats":
        ks = {KERNEL_NAMES[i]: {"launches": s.k[i].launches, "ms": s.k[i].ms, "bytes": s.k[i].bytes}
              for i in range(7)}
        vals = {f: getattr(s, f) for f, _ in _Stats._fields_ if f != "k"}
        return cls(kernels=ks, **vals)


SCHED_LEVELS = 64
ACT_TD, ACT_BU, ACT_STOP, ACT_POLICY = 0, 1, 2, -1


def schedule_array(fw=None, bw=None, cc=None) -> np.ndarray:
    """Forced-decision table for msc_opts.schedule: per search (FW, BW, CC) a list whose
    entry L-1 is the action after level L (ACT_TD / ACT_BU / ACT_STOP, ACT_POLICY = free)."""
    s = np.full(3 * SCHED_LEVELS, -1, dtype=np.int8)
    for k, seq in enumerate((fw, bw, cc)):
        if seq is not None:
            seq = list(seq)[:SCHED_LEVELS]
            s[k * SCHED_LEVELS:k * SCHED_LEVELS + len(seq)] = seq
    return s


def _opts(policy, labels, timing, record, validate, ctx, labels_where=HOST, alpha=15.0, beta=25,
          device=0, schedule=None, trim_level=1, replay=False) -> _Opts:
    o = _Opts()
    _lib().msc_opts_default(C.byref(o))
    if schedule is not None:
        o.schedule = schedule.ctypes.data
        o.schedule_len = int(schedule.shape[0])
    o.alpha, o.beta = alpha, beta
    o.trim_level = trim_level
    o.policy = policy
    o.label_mode = labels
    o.labels_where = labels_where
    o.timing = 1 if timing else 0
    o.record = 1 if record else 0
    o.validate = 1 if validate else 0
    o.ctx = ctx.handle if ctx is not None else None
    o.device = ctx.device if ctx is not None else device
    o.replay = 1 if replay else 0
    return o


def _run(fn_name, g, max_deg_vert, avg_degree, policy, labels, timing, record, validate, ctx, alpha, beta,
         out, schedule, batch=None, trim_level=1, fwbw_rounds=1, replay=False):
    n = g.n
    where = HOST
    if isinstance(out, DeviceLabels):
        if out.n != n:
            raise ValueError("DeviceLabels size differs from the graph")
        if ctx is None or out.ctx is not ctx:
            raise ValueError("DeviceLabels must belong to the context used for the run")
        where, dst = DEVICE, out.ptr.value
    else:
        if out is None:
            out = np.empty(n, dtype=np.int32)
        elif out.dtype != np.int32 or out.shape != (n,) or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous int32 array of length n")
        dst = _ptr(out) if n else None
    if schedule is not None:
        schedule = np.ascontiguousarray(schedule, dtype=np.int8)
    o = _opts(policy, labels, timing, record, validate, ctx, labels_where=where, alpha=alpha, beta=beta,
              schedule=schedule, trim_level=trim_level, replay=replay)
    o.fwbw_rounds = int(fwbw_rounds)
    st = _Stats()
    csr = g.csr()
    piv = -1 if max_deg_vert is None else int(max_deg_vert)
    avg = -1.0 if avg_degree is None else float(avg_degree)
    if batch is None:
        rc = getattr(_lib(), fn_name)(C.byref(csr), piv, avg, dst, C.byref(o), C.byref(st))
    else:
        base, pivots = batch
        rc = getattr(_lib(), fn_name)(C.byref(csr), int(base.shape[0] - 1), _ptr(base),
                                      _ptr(pivots) if pivots is not None else None, dst, C.byref(o), C.byref(st))
    _check(rc, fn_name)
    return out, Stats._from(st)


class Prepared:
    """One decomposition call with its arguments marshalled once (for timing loops): each
    call is a single C ABI call with no statistics requested.  ``kind``: "scc" or "cc";
    ``batch``: (vert_base, pivots) for the batched entry points; ``out``: DeviceLabels."""

    def __init__(self, kind, g, max_deg_vert=None, avg_degree=None, *, policy=POLICY_PLAIN,
                 labels=LABELS_CANONICAL, ctx: Context, out: "DeviceLabels", batch=None, replay=False):
        if out.n != g.n or out.ctx is not ctx:
            raise ValueError("out must be a DeviceLabels of the graph's size on the same context")
        self._keep = (g, out)
        self._csr = g.csr()
        self._opts = _opts(policy, labels, False, False, False, ctx, labels_where=DEVICE, replay=replay)
        self._dst = out.ptr.value
        piv = -1 if max_deg_vert is None else int(max_deg_vert)
        avg = -1.0 if avg_degree is None else float(avg_degree)
        if batch is None:
            fn = _lib().msc_scc_run if kind == "scc" else _lib().msc_cc_run
            self._args = (C.byref(self._csr), piv, avg, self._dst, C.byref(self._opts), None)
        else:
            base, pivots = batch
            self._base = np.ascontiguousarray(base, dtype=np.int64)
            self._piv = None if pivots is None else np.ascontiguousarray(pivots, dtype=np.int32)
            fn = _lib().msc_scc_run_batch if kind == "scc" else _lib().msc_cc_run_batch
            self._args = (C.byref(self._csr), int(self._base.shape[0] - 1), _ptr(self._base),
                          _ptr(self._piv) if self._piv is not None else None, self._dst, C.byref(self._opts), None)
        self._fn = fn
        self._name = fn.__name__

    def __call__(self) -> None:
        rc = self._fn(*self._args)
        if rc != 0:
            _check(rc, self._name)


def run_scc(g, max_deg_vert: int | None = None, avg_degree: float | None = None, *,
            policy: int = POLICY_PLAIN, labels: int = LABELS_CANONICAL, timing: bool = False,
            record: bool = False, validate: bool = False, ctx: Context | None = None,
            alpha: float = 15.0, beta: int = 25, out: np.ndarray | None = None, stats: bool = False,
            schedule: np.ndarray | None = None, trim_level: int = 1, fwbw_rounds: int = 1,
            replay: bool = False):
    """SCC labels of ``g`` (a Graph or DeviceGraph) on the GPU.

    run_scc (multistep/scc_main.cpp:210).  ``max_deg_vert`` / ``avg_degree`` default to
    create_csr's rule computed on device.  Returns labels, or (labels, Stats) with
    ``stats=True``.  ``schedule`` (see :func:`schedule_array`) forces per-level decisions.
    ``trim_level`` 2 is the reference's complete trim (TRIM_LEVEL 2, scc_trim.cpp:141-297).
    ``fwbw_rounds`` > 1 repeats FW-BW from new pivots while the SCC found is large
    (msc_opts.fwbw_rounds); the labels do not change.  ``replay`` lets the context capture the
    call's launches in a HIP graph and replay it on later identical calls (msc_opts.replay)."""
    lab, st = _run("msc_scc_run", g, max_deg_vert, avg_degree, policy, labels, timing, record, validate, ctx,
                   alpha, beta, out, schedule, trim_level=trim_level, fwbw_rounds=fwbw_rounds, replay=replay)
    return (lab, st) if stats else lab


def run_scc_noalloc(g: "DeviceGraph", max_deg_vert: int | None = None, avg_degree: float | None = None, *,
                    ctx: Context, out: "DeviceLabels", policy: int = POLICY_PLAIN,
                    labels: int = LABELS_CANONICAL, stats: bool = False):
    """run_scc through msc_scc_run_noalloc: the same labels, computed in the library's static
    per-device scratch array instead of a workspace it allocates (for callers that cannot
    allocate).  ``g`` and ``out`` must be device-resident and ``ctx`` created beforehand;
    n <= 2**20 and m <= 2**24.  Returns ``out`` (or (out, Stats))."""
    lab, st = _run("msc_scc_run_noalloc", g, max_deg_vert, avg_degree, policy, labels, False, False, False, ctx,
                   15.0, 25, out, None)
    return (lab, st) if stats else lab


def run_cc(g, max_deg_vert: int | None = None, avg_degree: float | None = None, *,
           policy: int = POLICY_PLAIN, labels: int = LABELS_CANONICAL, timing: bool = False,
           record: bool = False, validate: bool = False, ctx: Context | None = None,
           alpha: float = 15.0, beta: int = 25, out: np.ndarray | None = None, stats: bool = False,
           schedule: np.ndarray | None = None, replay: bool = False):
    """Weakly connected components (run_cc, multistep/cc_main.cpp:184, on the mirror)."""
    lab, st = _run("msc_cc_run", g, max_deg_vert, avg_degree, policy, labels, timing, record, validate, ctx,
                   alpha, beta, out, schedule, replay=replay)
    return (lab, st) if stats else lab


# ---- edge-list text on the device -----------------------------------------------------------
def format_edge_list(g: Graph) -> bytes:
    """The reference's edge-list text of g ("n m" then "u v" per edge, CSR order)."""
    hg = _HostGraph(g.n, g.m, g.out_off.ctypes.data_as(C.POINTER(C.c_uint32)),
                    g.out_adj.ctypes.data_as(C.POINTER(C.c_int32)), g.in_off.ctypes.data_as(C.POINTER(C.c_uint32)),
                    g.in_adj.ctypes.data_as(C.POINTER(C.c_int32)), g.max_deg_vert, g.avg_degree)
    size = C.c_int64()
    _check(_lib().msc_format_edge_list(C.byref(hg), None, 0, C.byref(size)), "msc_format_edge_list")
    buf = C.create_string_buffer(size.value + 16)
    _check(_lib().msc_format_edge_list(C.byref(hg), buf, size.value + 16, C.byref(size)), "msc_format_edge_list")
    return buf.raw[:size.value]


def _ingested(ctx: Context, call, what: str) -> DeviceGraph:
    dg = DeviceGraph(ctx, None)
    piv, avg = C.c_int32(), C.c_double()
    _check(call(C.byref(dg._csr), C.byref(piv), C.byref(avg)), what)
    dg.n, dg.m = dg._csr.n, dg._csr.m
    dg.max_deg_vert, dg.avg_degree = piv.value, avg.value
    return dg


def ingest_text(ctx: Context, text: bytes) -> DeviceGraph:
    """read_edge + create_csr on the device (msc_graph_ingest_text)."""
    return _ingested(ctx, lambda d, p, a: _lib().msc_graph_ingest_text(ctx.handle, text, len(text), d, p, a),
                     "msc_graph_ingest_text")


def read_edge_device(ctx: Context, path: str) -> DeviceGraph:
    """read_edge + create_csr on the device from a file (msc_graph_read_edge_list_device)."""
    return _ingested(ctx, lambda d, p, a: _lib().msc_graph_read_edge_list_device(ctx.handle, os.fsencode(path), d,
                                                                                  p, a),
                     f"msc_graph_read_edge_list_device({path})")


# ---- batches -------------------------------------------------------------------------------
def pack_graphs(graphs) -> tuple[Graph, np.ndarray]:
    """Packs graphs into one CSR (msc_batch_pack): returns (union, vert_base[len+1])."""
    k = len(graphs)
    arr = (C.POINTER(_HostGraph) * k)()
    keep = []
    for i, g in enumerate(graphs):
        hg = _HostGraph(g.n, g.m, g.out_off.ctypes.data_as(C.POINTER(C.c_uint32)),
                        g.out_adj.ctypes.data_as(C.POINTER(C.c_int32)),
                        g.in_off.ctypes.data_as(C.POINTER(C.c_uint32)),
                        g.in_adj.ctypes.data_as(C.POINTER(C.c_int32)), g.max_deg_vert, g.avg_degree)
        keep.append(hg)
        arr[i] = C.pointer(hg)
    base = np.empty(k + 1, dtype=np.int64)
    out = C.POINTER(_HostGraph)()
    _check(_lib().msc_batch_pack(k, arr, C.byref(out), _ptr(base)), "msc_batch_pack")
    return _wrap(out), base


def gen_batch(num_graphs: int = 8192, median_n: float = 4096, sigma: float = 1.4, min_n: int = 256,
              max_n: int = 500_000, seed: int = 1) -> tuple[Graph, np.ndarray, np.ndarray]:
    """Synthetic packed batch (msc_gen_batch): (union, vert_base, kinds; 0 R-MAT, 1 planted)."""
    base = np.empty(num_graphs + 1, dtype=np.int64)
    kinds = np.empty(num_graphs, dtype=np.int32)
    out = C.POINTER(_HostGraph)()
    _check(_lib().msc_gen_batch(num_graphs, median_n, sigma, min_n, max_n, seed, C.byref(out), _ptr(base),
                                _ptr(kinds)), "msc_gen_batch")
    return _wrap(out), base, kinds


def batch_pivots(g, vert_base) -> np.ndarray:
    """create_csr's pivot (multistep/scc_main.cpp:169-183) of every graph of a packed batch,
    in the graph's own ids: the first vertex with the strictly largest out x in degree product
    (as doubles), -1 when every product is 0.  These are the pivots run_scc_batch computes
    on device when ``pivots`` is None; computing them once with the batch is the
    counterpart of create_csr computing max_deg_vert with the graph."""
    base = np.ascontiguousarray(vert_base, dtype=np.int64)
    ng = base.shape[0] - 1
    piv = np.full(ng, -1, dtype=np.int32)
    prod = (np.diff(np.asarray(g.out_off, dtype=np.int64)).astype(np.float64)
            * np.diff(np.asarray(g.in_off, dtype=np.int64)).astype(np.float64))
    sizes = np.diff(base)
    live = np.flatnonzero(sizes > 0)
    if live.size == 0:
        return piv
    best = np.maximum.reduceat(prod, base[live])
    hit = np.flatnonzero((prod > 0.0) & (prod == np.repeat(best, sizes[live])))
    owner = np.searchsorted(base, hit, side="right") - 1
    first_graphs, first_at = np.unique(owner, return_index=True)
    piv[first_graphs] = (hit[first_at] - base[first_graphs]).astype(np.int32)
    return piv


def _batch_args(vert_base, pivots):
    base = np.ascontiguousarray(vert_base, dtype=np.int64)
    piv = None if pivots is None else np.ascontiguousarray(pivots, dtype=np.int32)
    if base.ndim != 1 or base.shape[0] < 2:
        raise ValueError("vert_base needs num_graphs + 1 entries")
    if piv is not None and piv.shape != (base.shape[0] - 1,):
        raise ValueError("pivots needs one entry per graph")
    return base, piv


def run_scc_batch(g, vert_base, pivots=None, *, policy: int = POLICY_PLAIN, labels: int = LABELS_CANONICAL,
                  timing: bool = False, record: bool = False, validate: bool = False, ctx: Context | None = None,
                  alpha: float = 15.0, beta: int = 25, out=None, stats: bool = False, trim_level: int = 1,
                  schedule: np.ndarray | None = None):
    """SCC labels of every graph of a packed batch (msc_scc_run_batch); each graph's slice
    holds its own local labels.  ``pivots`` (local ids, -1 none) default to create_csr's rule.
    ``schedule`` forces the actions of the batch's merged FW / BW searches (schedule_array)."""
    lab, st = _run("msc_scc_run_batch", g, None, None, policy, labels, timing, record, validate, ctx, alpha, beta,
                   out, schedule, batch=_batch_args(vert_base, pivots), trim_level=trim_level)
    return (lab, st) if stats else lab


def run_cc_batch(g, vert_base, pivots=None, *, policy: int = POLICY_PLAIN, labels: int = LABELS_CANONICAL,
                 timing: bool = False, record: bool = False, validate: bool = False, ctx: Context | None = None,
                 alpha: float = 15.0, beta: int = 25, out=None, stats: bool = False,
                 schedule: np.ndarray | None = None):
    """Weakly connected components of every graph of a packed batch (msc_cc_run_batch)."""
    lab, st = _run("msc_cc_run_batch", g, None, None, policy, labels, timing, record, validate, ctx, alpha, beta,
                   out, schedule, batch=_batch_args(vert_base, pivots))
    return (lab, st) if stats else lab


# ---- learned policy (MLP 12 -> 32 -> 32 -> 3) ------------------------------------------
POLICY_WEIGHTS = os.path.join(_HERE, "policy_weights.npz")


def set_policy(ctx: Context, blob: np.ndarray, blob_cc: np.ndarray | None = None) -> None:
    """Loads MLP weights + standardisation (MLP_BLOB float32) into the context: one network
    for every search, or with ``blob_cc`` a second one for the CC search."""
    b = np.ascontiguousarray(blob, dtype=np.float32)
    if b.shape != (MLP_BLOB,) or (blob_cc is not None and np.shape(blob_cc) != (MLP_BLOB,)):
        raise ValueError(f"a policy blob must hold {MLP_BLOB} floats")
    if blob_cc is not None:
        b = np.ascontiguousarray(np.concatenate([b, np.asarray(blob_cc, dtype=np.float32)]))
    _check(_lib().msc_policy_set_weights(ctx.handle, _ptr(b), b.shape[0]), "msc_policy_set_weights")


def load_policy(ctx: Context, path: str = POLICY_WEIGHTS) -> np.ndarray:
    """Loads the shipped (or given) trained weights into ctx (`blob`, and `blob_cc` for the CC
    search when the file has one); returns the first blob."""
    with np.load(path, allow_pickle=False) as z:
        blob = z["blob"].astype(np.float32)
        blob_cc = z["blob_cc"].astype(np.float32) if "blob_cc" in z.files else None
    set_policy(ctx, blob, blob_cc)
    return blob


def mlp_forward(ctx: Context, blob: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Logits [count, 3] of raw features x [count, 12], computed by the device kernel."""
    b = np.ascontiguousarray(blob, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
    out = np.empty((xx.shape[0], MLP_OUT), dtype=np.float32)
    _check(_lib().msc_mlp_forward(ctx.handle, _ptr(b), _ptr(xx), xx.shape[0], _ptr(out)), "msc_mlp_forward")
    return out


def mlp_grad(ctx: Context, blob, x, y, w=None) -> tuple[float, np.ndarray]:
    """(sum_b w_b CE_b, its gradient w.r.t. the MLP_PARAMS weights) from the device kernels."""
    b = np.ascontiguousarray(blob, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
    yy = np.ascontiguousarray(y, dtype=np.int32)
    ww = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
    grad = np.empty(MLP_PARAMS, dtype=np.float32)
    loss = C.c_float()
    _check(_lib().msc_mlp_grad(ctx.handle, _ptr(b), _ptr(xx), _ptr(yy), None if ww is None else _ptr(ww),
                               xx.shape[0], C.byref(loss), _ptr(grad)), "msc_mlp_grad")
    return float(loss.value), grad


def mlp_train(ctx: Context, blob, x, y, w=None, *, epochs=30, batch=1024, lr=3e-3, beta1=0.9, beta2=0.999,
              eps=1e-8, weight_decay=0.0, seed=1) -> tuple[np.ndarray, np.ndarray]:
    """Adam on device; returns (trained blob, per-epoch mean loss)."""
    b = np.array(blob, dtype=np.float32, copy=True)
    xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
    yy = np.ascontiguousarray(y, dtype=np.int32)
    ww = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
    curve = np.zeros(epochs, dtype=np.float32)
    o = _TrainOpts(epochs, batch, lr, beta1, beta2, eps, weight_decay, seed)
    _check(_lib().msc_mlp_train(ctx.handle, _ptr(b), _ptr(xx), _ptr(yy), None if ww is None else _ptr(ww),
                                xx.shape[0], C.byref(o), _ptr(curve)), "msc_mlp_train")
    return b, curve


def scc_pivot(g, ctx: Context | None = None) -> tuple[int, float]:
    """create_csr's pivot and average degree, computed on device."""
    o = _opts(POLICY_PLAIN, LABELS_CANONICAL, False, False, False, ctx)
    piv, avg = C.c_int32(), C.c_double()
    csr = g.csr()
    _check(_lib().msc_scc_pivot(C.byref(csr), C.byref(o), C.byref(piv), C.byref(avg)), "msc_scc_pivot")
    return piv.value, avg.value


def cc_pivot(g, ctx: Context | None = None) -> tuple[int, float]:
    """cc_main's create_csr pivot (largest degree in the mirrored graph) and average degree
    (2m/n), computed on device: the max_deg_vert / avg_degree run_cc takes."""
    o = _opts(POLICY_PLAIN, LABELS_CANONICAL, False, False, False, ctx)
    piv, avg = C.c_int32(), C.c_double()
    csr = g.csr()
    _check(_lib().msc_cc_pivot(C.byref(csr), C.byref(o), C.byref(piv), C.byref(avg)), "msc_cc_pivot")
    return piv.value, avg.value


def _verify(fn, labels):
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    v = _Verify()
    _check(getattr(_lib(), fn)(lab.shape[0], _ptr(lab) if lab.size else None, C.byref(v)), fn)
    return v


def scc_verify(labels) -> tuple[int, int, int, int]:
    """(num SCCs, nontrivial, max SCC, unassigned) -- multistep/scc_verify.cpp:47-77."""
    v = _verify("msc_scc_verify", labels)
    return v.num_components, v.nontrivial, v.max_size, v.unassigned


def cc_verify(labels) -> tuple[int, int, int]:
    """(num CCs, max CC, unassigned) -- multistep/cc_verify.cpp:47-74."""
    v = _verify("msc_cc_verify", labels)
    return v.num_components, v.max_size, v.unassigned


def verify_line(labels, kind: str = "scc") -> str:
    """The reference's summary line for these labels."""
    v = _verify("msc_scc_verify" if kind == "scc" else "msc_cc_verify", labels)
    buf = C.create_string_buffer(256)
    _check(_lib().msc_verify_format(C.byref(v), 0 if kind == "scc" else 1, buf, 256), "msc_verify_format")
    return buf.value.decode()


def canonicalize(labels) -> np.ndarray:
    """Copy of ``labels`` with every component relabelled to its minimum vertex id."""
    lab = np.array(labels, dtype=np.int32, copy=True)
    _check(_lib().msc_canonicalize(lab.shape[0], _ptr(lab) if lab.size else None), "msc_canonicalize")
    return lab


def write_labels(path: str, labels) -> None:
    """output_scc (multistep/scc_main.cpp:193-202)."""
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    _check(_lib().msc_write_labels(os.fsencode(path), lab.shape[0], _ptr(lab) if lab.size else None),
           "write_labels")
