"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and, when it has been built,
to the compiled reference harness (oracle/_ref/libvite_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the bench's
CPU-baseline leg as the checker.  Nothing in the product imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libvite_ref.so")
REF_TRACE = os.path.join(HERE, "_ref", "ref_trace")
REF_BIN = os.path.join(HERE, "_ref", "graphClustering")

EDGE_DTYPE = np.dtype([("tail", "<i8"), ("weight", "<f8")])

i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)


class _Graph(C.Structure):
    _fields_ = [("nv", C.c_int64), ("ne", C.c_int64), ("offs", i64p), ("edges", C.c_void_p)]


class _PhaseParams(C.Structure):
    _fields_ = [("lower", C.c_double), ("thresh", C.c_double), ("et_type", C.c_int),
                ("et_delta", C.c_double), ("mode", C.c_int), ("num_color", C.c_int64),
                ("colors", i64p)]


class _Trace(C.Structure):
    _fields_ = [("max_iters", C.c_int), ("n", C.c_int), ("q", f64p), ("moves", i64p)]


class _RunParams(C.Structure):
    _fields_ = [("threshold", C.c_double), ("threshold_cycling", C.c_int), ("et_type", C.c_int),
                ("et_delta", C.c_double), ("coloring", C.c_int), ("vertex_order", C.c_int),
                ("max_colors", C.c_int), ("single_color_iter", C.c_int), ("one_phase", C.c_int)]


MAXL = 256


class _RunResult(C.Structure):
    _fields_ = [("nlevels", C.c_int), ("level_q", C.c_double * MAXL), ("level_iters", C.c_int * MAXL),
                ("level_nv", C.c_int64 * MAXL), ("level_ne", C.c_int64 * MAXL), ("ncalls", C.c_int),
                ("call_q", C.c_double * (MAXL + 1)), ("call_iters", C.c_int * (MAXL + 1)),
                ("call_cvect_fnv", C.c_uint64 * (MAXL + 1)),
                ("final_q", C.c_double), ("total_iters", C.c_int), ("num_colors", C.c_int64)]


@dataclass
class Graph:
    """CSR graph in Vite's layout: offsets int64[nv+1], edges {int64 tail, float64 weight}[ne]."""
    offs: np.ndarray
    edges: np.ndarray

    @property
    def nv(self) -> int:
        return int(self.offs.shape[0] - 1)

    @property
    def ne(self) -> int:
        return int(self.edges.shape[0])

    def rows(self):
        return self.offs, self.edges["tail"], self.edges["weight"]


def _lib():
    if not hasattr(_lib, "h"):
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        h = C.CDLL(LIB_PATH)
        h.or_louvain_phase.restype = C.c_double
        h.or_init_degrees.restype = C.c_double
        h.or_modularity.restype = C.c_double
        h.or_color.restype = C.c_int64
        h.or_rgg_radius.restype = C.c_double
        h.or_rng_u64.restype = C.c_uint64
        h.or_rng_u64.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        h.or_color_hash.restype = C.c_uint32
        h.or_color_hash.argtypes = [C.c_uint32, C.c_uint32]
        h.or_seed_seq_1.restype = C.c_uint32
        h.or_seed_seq_1.argtypes = [C.c_uint32]
        h.or_fnv_i64.restype = C.c_uint64
        h.or_gen_rgg.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.POINTER(_Graph)]
        h.or_gen_ba.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.POINTER(_Graph)]
        h.or_gen_vite_rgg.argtypes = [C.c_int64, C.POINTER(_Graph)]
        _lib.h = h
    return _lib.h


def ref_available() -> bool:
    return os.path.exists(REF_LIB_PATH) and os.path.exists(REF_TRACE)


def _ref():
    if not hasattr(_ref, "h"):
        h = C.CDLL(REF_LIB_PATH)
        h.ref_phase.restype = C.c_double
        h.ref_init_degrees.restype = C.c_double
        h.ref_modularity.restype = C.c_double
        h.ref_color.restype = C.c_long
        h.ref_sweep_omp.restype = C.c_double
        _ref.h = h
    return _ref.h


def _p(a, t=i64p):
    return a.ctypes.data_as(t)


def _cg(g: Graph) -> _Graph:
    return _Graph(g.nv, g.ne, _p(g.offs), g.edges.ctypes.data_as(C.c_void_p))


def _take(cg: _Graph) -> Graph:
    nv, ne = cg.nv, cg.ne
    offs = np.ctypeslib.as_array(cg.offs, shape=(nv + 1,)).copy()
    if ne:
        buf = (C.c_char * (16 * ne)).from_address(cg.edges)
        edges = np.frombuffer(bytes(buf), dtype=EDGE_DTYPE).copy()
    else:
        edges = np.zeros(0, dtype=EDGE_DTYPE)
    _lib().or_graph_free(C.byref(cg))
    return Graph(offs, edges)


def make_graph(offs, tails, weights) -> Graph:
    offs = np.ascontiguousarray(offs, dtype=np.int64)
    e = np.zeros(len(tails), dtype=EDGE_DTYPE)
    e["tail"] = tails
    e["weight"] = weights
    return Graph(offs, e)


def from_edge_list(nv: int, pairs, weights=None, directed=False) -> Graph:
    """CSR from (u, v[, w]) pairs; undirected pairs are mirrored; rows sorted by tail."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    w = np.ones(len(pairs)) if weights is None else np.asarray(weights, dtype=np.float64)
    if not directed:
        loops = pairs[:, 0] == pairs[:, 1]
        pairs = np.concatenate([pairs, pairs[~loops][:, ::-1]])
        w = np.concatenate([w, w[~loops]])
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    pairs, w = pairs[order], w[order]
    offs = np.zeros(nv + 1, dtype=np.int64)
    np.add.at(offs, pairs[:, 0] + 1, 1)
    offs = np.cumsum(offs)
    return make_graph(offs, pairs[:, 1], w)


# ---------------------------------------------------------------------------------
# Oracle calls
# ---------------------------------------------------------------------------------

def gen_rgg(seed: int, n: int, extra_pct: float = 0.0) -> Graph:
    cg = _Graph()
    rc = _lib().or_gen_rgg(seed, n, extra_pct, C.byref(cg))
    assert rc == 0, rc
    return _take(cg)


def gen_ba(seed: int, n: int, m: int) -> Graph:
    cg = _Graph()
    rc = _lib().or_gen_ba(seed, n, m, C.byref(cg))
    assert rc == 0, rc
    return _take(cg)


def gen_vite_rgg(n: int) -> Graph:
    cg = _Graph()
    rc = _lib().or_gen_vite_rgg(n, C.byref(cg))
    assert rc == 0, rc
    return _take(cg)


def rgg_radius(n: int) -> float:
    return _lib().or_rgg_radius(C.c_int64(n))


def init_degrees(g: Graph):
    vdeg = np.zeros(g.nv)
    c = _lib().or_init_degrees(C.byref(_cg(g)), _p(vdeg, f64p))
    return vdeg, c


def sweep(g: Graph, curr, csize, cdeg, vdeg, constant, order=None):
    nv = g.nv
    curr = np.ascontiguousarray(curr, np.int64)
    csize = np.ascontiguousarray(csize, np.int64)
    cdeg = np.ascontiguousarray(cdeg, np.float64)
    vdeg = np.ascontiguousarray(vdeg, np.float64)
    if order is None:
        order = np.arange(nv, dtype=np.int64)
    order = np.ascontiguousarray(order, np.int64)
    target = np.zeros(nv, np.int64)
    cw = np.zeros(nv)
    dsize = np.zeros(nv, np.int64)
    ddeg = np.zeros(nv)
    _lib().or_sweep(C.byref(_cg(g)), _p(curr), _p(csize), _p(cdeg, f64p), _p(vdeg, f64p),
                    C.c_double(constant), _p(order), C.c_int64(len(order)), _p(target),
                    _p(cw, f64p), _p(dsize), _p(ddeg, f64p))
    return target, cw, dsize, ddeg


def modularity(cw, cdeg, constant) -> float:
    cw = np.ascontiguousarray(cw, np.float64)
    cdeg = np.ascontiguousarray(cdeg, np.float64)
    return _lib().or_modularity(C.c_int64(len(cw)), _p(cw, f64p), _p(cdeg, f64p), C.c_double(constant))


MODE = {"plain": 0, "color": 1, "vorder": 2}


def louvain_phase(g: Graph, lower=-1.0, thresh=1e-6, et_type=0, et_delta=1.0, mode="plain",
                  num_color=0, colors=None, trace_cap=0):
    p = _PhaseParams(lower, thresh, et_type, et_delta, MODE[mode], num_color, None)
    if colors is not None:
        colors = np.ascontiguousarray(colors, np.int64)
        p.colors = _p(colors)
    cvect = np.zeros(g.nv, np.int64)
    iters = C.c_int(0)
    tr = None
    if trace_cap:
        tq = np.zeros(trace_cap)
        tm = np.zeros(trace_cap, np.int64)
        tr = _Trace(trace_cap, 0, _p(tq, f64p), _p(tm))
    q = _lib().or_louvain_phase(C.byref(_cg(g)), C.byref(p), _p(cvect), C.byref(iters),
                                C.byref(tr) if tr is not None else None)
    if trace_cap:
        return q, cvect, iters.value, tq[:tr.n].copy(), tm[:tr.n].copy()
    return q, cvect, iters.value


def rebuild(g: Graph, cvect) -> Graph:
    cvect = np.ascontiguousarray(cvect, np.int64)
    out = _Graph()
    rc = _lib().or_rebuild(C.byref(_cg(g)), _p(cvect), C.byref(out))
    assert rc == 0, rc
    return _take(out)


def color(g: Graph, nhash: int, target_pct: int = 70, single: bool = False):
    colors = np.zeros(g.nv, np.int64)
    nc = _lib().or_color(C.byref(_cg(g)), C.c_int64(nhash), C.c_int(target_pct), C.c_int(int(single)), _p(colors))
    return nc, colors


def color_hash(a: int, seed: int) -> int:
    return _lib().or_color_hash(a & 0xffffffff, seed & 0xffffffff)


@dataclass
class RunResult:
    level_q: list
    level_iters: list
    level_nv: list
    level_ne: list
    call_q: list
    call_iters: list
    final_q: float
    total_iters: int
    num_colors: int
    call_cvect_fnv: list


def run(g: Graph, threshold=1e-6, threshold_cycling=False, et_type=0, et_delta=1.0,
        coloring=False, vertex_order=False, max_colors=8, single_color_iter=False,
        one_phase=False) -> RunResult:
    rp = _RunParams(threshold, int(threshold_cycling), et_type, et_delta, int(coloring),
                    int(vertex_order), max_colors, int(single_color_iter), int(one_phase))
    res = _RunResult()
    rc = _lib().or_run(C.byref(_cg(g)), C.byref(rp), C.byref(res))
    assert rc == 0, rc
    n, m = res.nlevels, res.ncalls
    return RunResult(list(res.level_q[:n]), list(res.level_iters[:n]), list(res.level_nv[:n]),
                     list(res.level_ne[:n]), list(res.call_q[:m]), list(res.call_iters[:m]),
                     res.final_q, res.total_iters, res.num_colors, [int(x) for x in res.call_cvect_fnv[:m]])


class _CompareStats(C.Structure):
    _fields_ = [("n", C.c_int64), ("nc_truth", C.c_int64), ("nc_found", C.c_int64), ("same_same", C.c_int64),
                ("same_diff", C.c_int64), ("diff_same", C.c_int64), ("precision", C.c_double),
                ("recall", C.c_double), ("f_score", C.c_double), ("gini_truth", C.c_double),
                ("gini_found", C.c_double)]


def compare(truth, found) -> dict:
    """compare_communities + compute_gini_coeff (compare.cpp:58-337) at one thread."""
    t = np.ascontiguousarray(truth, np.int64)
    f = np.ascontiguousarray(found, np.int64)
    assert len(t) == len(f)
    st = _CompareStats()
    rc = _lib().or_compare(C.c_int64(len(t)), _p(t), _p(f), C.byref(st))
    if rc != 0:
        raise ValueError(f"or_compare: {rc}")
    return {k: getattr(st, k) for k, _ in _CompareStats._fields_}


def fnv_i64(v) -> int:
    v = np.ascontiguousarray(v, np.int64)
    return int(_lib().or_fnv_i64(_p(v), C.c_int64(len(v))))


def write_bin(path: str, g: Graph) -> None:
    with open(path, "wb") as f:
        np.array([g.nv, g.ne], dtype=np.int64).tofile(f)
        g.offs.astype(np.int64).tofile(f)
        g.edges.tofile(f)


def read_bin(path: str) -> Graph:
    with open(path, "rb") as f:
        nv, ne = np.fromfile(f, dtype=np.int64, count=2)
        offs = np.fromfile(f, dtype=np.int64, count=int(nv) + 1)
        edges = np.fromfile(f, dtype=EDGE_DTYPE, count=int(ne))
    return Graph(offs - offs[0], edges)


# ---------------------------------------------------------------------------------
# Compiled-reference calls (only where oracle/_ref was built)
# ---------------------------------------------------------------------------------

def ref_sweep(g: Graph, curr, csize, cdeg, vdeg, constant, order=None):
    nv = g.nv
    curr = np.ascontiguousarray(curr, np.int64)
    csize = np.ascontiguousarray(csize, np.int64)
    cdeg = np.ascontiguousarray(cdeg, np.float64)
    vdeg = np.ascontiguousarray(vdeg, np.float64)
    if order is None:
        order = np.arange(nv, dtype=np.int64)
    order = np.ascontiguousarray(order, np.int64)
    target = np.zeros(nv, np.int64)
    cw = np.zeros(nv)
    dsize = np.zeros(nv, np.int64)
    ddeg = np.zeros(nv)
    _ref().ref_sweep(C.c_int64(nv), C.c_int64(g.ne), _p(g.offs), g.edges.ctypes.data_as(C.c_void_p),
                     _p(curr), _p(csize), _p(cdeg, f64p), _p(vdeg, f64p), C.c_double(constant),
                     _p(order), C.c_int64(len(order)), _p(target), _p(cw, f64p), _p(dsize), _p(ddeg, f64p))
    # vertices outside `order` keep the reference's untouched target (-1); report curr there
    mask = np.ones(nv, bool)
    mask[order] = False
    target[mask] = curr[mask]
    return target, cw, dsize, ddeg


def ref_sweep_omp(g: Graph, curr, csize, cdeg, vdeg, constant, v0: int, v1: int, threads: int, reps: int = 1):
    """The reference's parallel sweep loop on vertices [v0, v1): (mean seconds, targets)."""
    nv = g.nv
    curr = np.ascontiguousarray(curr, np.int64)
    csize = np.ascontiguousarray(csize, np.int64)
    cdeg = np.ascontiguousarray(cdeg, np.float64)
    vdeg = np.ascontiguousarray(vdeg, np.float64)
    target = np.zeros(nv, np.int64)
    sec = _ref().ref_sweep_omp(C.c_int64(nv), C.c_int64(g.ne), _p(g.offs), g.edges.ctypes.data_as(C.c_void_p),
                               _p(curr), _p(csize), _p(cdeg, f64p), _p(vdeg, f64p), C.c_double(constant),
                               C.c_int64(v0), C.c_int64(v1), C.c_int(threads), C.c_int(reps), _p(target))
    return sec, target


def ref_init_degrees(g: Graph):
    vdeg = np.zeros(g.nv)
    c = _ref().ref_init_degrees(C.c_int64(g.nv), C.c_int64(g.ne), _p(g.offs),
                                g.edges.ctypes.data_as(C.c_void_p), _p(vdeg, f64p))
    return vdeg, c


def ref_phase(g: Graph, lower=-1.0, thresh=1e-6, et_type=0, et_delta=1.0, mode="plain",
              num_color=0, colors=None):
    cvect = np.zeros(g.nv, np.int64)
    iters = C.c_int(0)
    cp = None
    if colors is not None:
        colors = np.ascontiguousarray(colors, np.int64)
        cp = _p(colors)
    q = _ref().ref_phase(C.c_int64(g.nv), C.c_int64(g.ne), _p(g.offs), g.edges.ctypes.data_as(C.c_void_p),
                         C.c_int(MODE[mode]), C.c_int(et_type), C.c_double(et_delta), C.c_long(num_color),
                         cp, C.c_double(lower), C.c_double(thresh), _p(cvect), C.byref(iters))
    return q, cvect, iters.value


def ref_rebuild(g: Graph, cvect) -> Graph:
    cvect = np.ascontiguousarray(cvect, np.int64)
    nv2, ne2 = C.c_int64(), C.c_int64()
    po, pe = i64p(), C.c_void_p()
    rc = _ref().ref_rebuild(C.c_int64(g.nv), C.c_int64(g.ne), _p(g.offs), g.edges.ctypes.data_as(C.c_void_p),
                            _p(cvect), C.byref(nv2), C.byref(ne2), C.byref(po), C.byref(pe))
    assert rc == 0
    offs = np.ctypeslib.as_array(po, shape=(nv2.value + 1,)).copy()
    edges = np.frombuffer(bytes((C.c_char * (16 * max(ne2.value, 1))).from_address(pe.value)),
                          dtype=EDGE_DTYPE)[:ne2.value].copy()
    _ref().ref_free(po)
    _ref().ref_free(pe)
    return Graph(offs, edges)


def ref_color(g: Graph, nhash: int, target_pct=70, single=False):
    colors = np.zeros(g.nv, np.int64)
    nc = _ref().ref_color(C.c_int64(g.nv), C.c_int64(g.ne), _p(g.offs), g.edges.ctypes.data_as(C.c_void_p),
                          C.c_long(nhash), C.c_int(target_pct), C.c_int(int(single)), _p(colors))
    return nc, colors


def ref_gen_rgg(n: int) -> Graph:
    nv2, ne2 = C.c_int64(), C.c_int64()
    po, pe = i64p(), C.c_void_p()
    _ref().ref_gen_rgg(C.c_int64(n), C.byref(nv2), C.byref(ne2), C.byref(po), C.byref(pe))
    offs = np.ctypeslib.as_array(po, shape=(nv2.value + 1,)).copy()
    edges = np.frombuffer(bytes((C.c_char * (16 * max(ne2.value, 1))).from_address(pe.value)),
                          dtype=EDGE_DTYPE)[:ne2.value].copy()
    _ref().ref_free(po)
    _ref().ref_free(pe)
    return Graph(offs, edges)


def ref_trace_run(bin_path: str, flags=(), threads=1, timeout=600):
    """Run the reference's own main() on a .bin file; returns the parsed TRACE records."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([REF_TRACE, "-f", bin_path, *flags], capture_output=True, text=True,
                         env=env, timeout=timeout, cwd=os.path.dirname(bin_path) or ".")
    if out.returncode != 0:
        raise RuntimeError(f"ref_trace failed: {out.stderr[-2000:]}")
    recs = []
    for line in out.stdout.splitlines():
        if not line.startswith("TRACE "):
            continue
        parts = line.split()
        rec = {"kind": parts[1]}
        for kv in parts[2:]:
            k, v = kv.split("=")
            rec[k] = v
        recs.append(rec)
    return recs, out.stdout
