"""vite_amd: Louvain community detection on one MI355X (gfx950).

Python view of the C ABI in include/vite_amd.h (ctypes over vite_amd/libvite_amd.so).
The shared library is required: importing the package raises if it has not been built,
there is no CPU fallback.

Graphs use Vite's CSR layout (graph.hpp:85-86): offsets int64[nv+1] and 16-byte edge
records {int64 tail, float64 weight} (edge.hpp:72-80).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvite_amd.so")

EDGE_DTYPE = np.dtype([("tail", "<i8"), ("weight", "<f8")])

VT_OK = 0
VT_ERR_ARG = -1
VT_ERR_HIP = -2
VT_ERR_BUDGET = -3
VT_ERR_LIMIT = -4
VT_ERR_STATE = -5

VT_MODE_PLAIN, VT_MODE_COLOR, VT_MODE_VORDER = 0, 1, 2
VT_FLAG_SEQUENTIAL_SUMS = 1
VT_FLAG_NO_GRAPH = 2
MAX_LEVELS = 256
# First word std::seed_seq{1} generates: the LCG start of Vite's generator (utils.hpp:163-170).
VITE_LCG_X0 = 1967017404

_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)


class _CompareStats(C.Structure):
    _fields_ = [("n", C.c_int64), ("nc_truth", C.c_int64), ("nc_found", C.c_int64), ("same_same", C.c_int64),
                ("same_diff", C.c_int64), ("diff_same", C.c_int64), ("precision", C.c_double),
                ("recall", C.c_double), ("f_score", C.c_double), ("gini_truth", C.c_double),
                ("gini_found", C.c_double)]


class VtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"vite_amd error {code}: {msg}")
        self.code = code


class _PhaseParams(C.Structure):
    _fields_ = [("lower", C.c_double), ("thresh", C.c_double), ("et_type", C.c_int32), ("mode", C.c_int32),
                ("et_delta", C.c_double), ("num_color", C.c_int64), ("colors", _i64p), ("flags", C.c_int32),
                ("max_iters", C.c_int32)]


class _RunParams(C.Structure):
    _fields_ = [("threshold", C.c_double), ("threshold_cycling", C.c_int32), ("et_type", C.c_int32),
                ("et_delta", C.c_double), ("coloring", C.c_int32), ("vertex_order", C.c_int32),
                ("max_colors", C.c_int32), ("single_color_iter", C.c_int32), ("one_phase", C.c_int32),
                ("flags", C.c_int32)]


class _RunResult(C.Structure):
    _fields_ = [("nlevels", C.c_int32), ("level_q", C.c_double * MAX_LEVELS),
                ("level_iters", C.c_int32 * MAX_LEVELS), ("level_nv", C.c_int64 * MAX_LEVELS),
                ("level_ne", C.c_int64 * MAX_LEVELS), ("ncalls", C.c_int32),
                ("call_q", C.c_double * (MAX_LEVELS + 1)), ("call_iters", C.c_int32 * (MAX_LEVELS + 1)),
                ("final_q", C.c_double), ("total_iters", C.c_int32), ("num_colors", C.c_int64),
                ("clustering_seconds", C.c_double), ("total_seconds", C.c_double)]


# Every symbol include/vite_amd.h declares, with its ctypes signature.
class _Tuning(C.Structure):
    _fields_ = [("bucket_max_degree", C.c_int32 * 6), ("blocks_per_bucket", C.c_int32),
                ("wide_table_max_degree", C.c_int32 * 2), ("vertex_records", C.c_int32),
                ("hub_block_min", C.c_int32)]


def _tuning_dict(t) -> dict:
    return dict(bucket_max_degree=list(t.bucket_max_degree), blocks_per_bucket=t.blocks_per_bucket,
                wide_table_max_degree=list(t.wide_table_max_degree), vertex_records=t.vertex_records,
                hub_block_min=t.hub_block_min)


class _Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("seed", C.c_uint64), ("n", C.c_int64), ("param", C.c_double),
                ("run", _RunParams)]


class _RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("slot", C.c_int32), ("submit_s", C.c_double), ("start_s", C.c_double),
                ("end_s", C.c_double), ("reserved_bytes", C.c_uint64), ("peak_bytes", C.c_uint64),
                ("nv", C.c_int64), ("ne", C.c_int64), ("batch", C.c_int32)]


class _BatchStats(C.Structure):
    _fields_ = [("phases", C.c_int32), ("rounds", C.c_int32), ("seconds", C.c_double), ("relists", C.c_int32)]


_SIGS = {
    "vt_ctx_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "vt_ctx_destroy": (C.c_int, [C.c_void_p]),
    "vt_last_error": (C.c_char_p, [C.c_void_p]),
    "vt_ctx_bytes_in_use": (C.c_uint64, [C.c_void_p]),
    "vt_ctx_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "vt_ctx_kernel_stats": (C.c_int, [C.c_void_p, C.c_char_p, _f64p, _i64p, _f64p]),
    "vt_ctx_reset_stats": (C.c_int, [C.c_void_p]),
    "vt_graph_upload": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, _i64p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "vt_graph_info": (C.c_int, [C.c_void_p, _i64p, _i64p]),
    "vt_graph_download": (C.c_int, [C.c_void_p, C.c_void_p, _i64p, C.c_void_p]),
    "vt_graph_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vt_louvain_phase": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_PhaseParams), _i64p,
                                   C.POINTER(C.c_int32), _f64p]),
    "vt_last_phase_trace": (C.c_int, [C.c_void_p, C.c_int32, _f64p, _i64p, C.POINTER(C.c_int32)]),
    "vt_phase_begin": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_PhaseParams), C.POINTER(C.c_void_p)]),
    "vt_phase_step": (C.c_int, [C.c_void_p, C.c_void_p, _f64p, C.POINTER(C.c_int32)]),
    "vt_phase_sweep": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vt_phase_end": (C.c_int, [C.c_void_p, C.c_void_p, _i64p, C.POINTER(C.c_int32), _f64p]),
    "vt_phase_get_state": (C.c_int, [C.c_void_p, C.c_void_p, _i64p, _i64p, _f64p, _f64p, _f64p, _i64p, _f64p]),
    "vt_sweep": (C.c_int, [C.c_void_p, C.c_void_p, _i64p, _i64p, _f64p, _f64p, C.c_double, _i64p, _f64p,
                           _i64p, _f64p]),
    "vt_modularity": (C.c_int, [C.c_void_p, C.c_int64, _f64p, _f64p, C.c_double, C.c_int32, _f64p]),
    "vt_modularity_noalloc": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_double, C.c_int32,
                                        C.c_void_p]),
    "vt_dev_sort_pairs": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p]),
    "vt_dev_scan": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]),
    "vt_init_degrees": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, _f64p, _f64p]),
    "vt_rebuild": (C.c_int, [C.c_void_p, C.c_void_p, _i64p, C.POINTER(C.c_void_p)]),
    "vt_color": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, _i64p, _i64p]),
    "vt_compare_communities": (C.c_int, [C.c_void_p, C.c_int64, _i64p, _i64p, C.c_void_p]),
    "vt_gen_rgg": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int64, C.c_double, C.POINTER(C.c_void_p)]),
    "vt_gen_ba": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, C.POINTER(C.c_void_p)]),
    "vt_gen_vite_rgg": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint32, C.POINTER(C.c_void_p)]),
    "vt_run": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_RunParams), C.POINTER(_RunResult)]),
    "vt_batch_create": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "vt_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "vt_batch_run": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_RunParams), C.POINTER(_RunResult),
                               C.POINTER(_BatchStats)]),
    "vt_batch_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.POINTER(_Tuning)]),
    "vt_batch_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vt_request_reservation": (C.c_uint64, [C.POINTER(_Request)]),
    "vt_server_create": (C.c_int, [C.c_int, C.c_uint64, C.c_int32, C.POINTER(C.c_void_p)]),
    "vt_server_submit": (C.c_int, [C.c_void_p, C.POINTER(_Request), C.POINTER(C.c_int64)]),
    "vt_server_set_batching": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64]),
    "vt_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(_RunResult), C.POINTER(_RequestInfo)]),
    "vt_server_last_error": (C.c_char_p, [C.c_void_p]),
    "vt_server_destroy": (C.c_int, [C.c_void_p]),
    "vt_ctx_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(_Tuning)]),
    "vt_ctx_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(_Tuning)]),
    "vt_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.POINTER(_Tuning)]),
}


def load_library(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise RuntimeError(f"vite_amd: native library not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = load_library()


def exported_symbols():
    return sorted(_SIGS)


def _p(a, t=_i64p):
    return a.ctypes.data_as(t)


@dataclass
class CSR:
    """Host CSR in Vite's layout."""
    offs: np.ndarray
    edges: np.ndarray

    @property
    def nv(self) -> int:
        return int(self.offs.shape[0] - 1)

    @property
    def ne(self) -> int:
        return int(self.edges.shape[0])

    @staticmethod
    def from_arrays(offs, tails, weights) -> "CSR":
        e = np.zeros(len(tails), dtype=EDGE_DTYPE)
        e["tail"] = tails
        e["weight"] = weights
        return CSR(np.ascontiguousarray(offs, dtype=np.int64), e)


class Context:
    """One device, one HIP stream, an optional HBM byte budget."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = C.c_void_p()
        rc = _lib.vt_ctx_create(device, hbm_budget, C.byref(h))
        if rc != VT_OK:
            raise VtError(rc, "vt_ctx_create failed (no GPU visible?)")
        self.h = h

    def close(self):
        if self.h:
            _lib.vt_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int):
        if rc != VT_OK:
            msg = _lib.vt_last_error(self.h)
            raise VtError(rc, msg.decode() if msg else "")

    @property
    def bytes_in_use(self) -> int:
        return int(_lib.vt_ctx_bytes_in_use(self.h))

    @property
    def tuning(self) -> dict:
        t = _Tuning()
        self.check(_lib.vt_ctx_get_tuning(self.h, C.byref(t)))
        return _tuning_dict(t)

    def set_tuning(self, bucket_max_degree=None, blocks_per_bucket=None, wide_table_max_degree=None,
                   vertex_records=None, hub_block_min=None):
        cur = self.tuning
        t = _Tuning((C.c_int32 * 6)(*(bucket_max_degree or cur["bucket_max_degree"])),
                    blocks_per_bucket or cur["blocks_per_bucket"],
                    (C.c_int32 * 2)(*(wide_table_max_degree or cur["wide_table_max_degree"])),
                    cur["vertex_records"] if vertex_records is None else int(vertex_records),
                    cur["hub_block_min"] if hub_block_min is None else int(hub_block_min))
        self.check(_lib.vt_ctx_set_tuning(self.h, C.byref(t)))

    def autotune(self, g: "DeviceGraph", prep: int = 2, reps: int = 3) -> dict:
        t = _Tuning()
        self.check(_lib.vt_autotune(self.h, g.h, prep, reps, C.byref(t)))
        return _tuning_dict(t)

    def set_profiling(self, on: bool):
        self.check(_lib.vt_ctx_set_profiling(self.h, int(on)))

    def kernel_stats(self, name: str):
        avg, cnt, tot = C.c_double(), C.c_int64(), C.c_double()
        self.check(_lib.vt_ctx_kernel_stats(self.h, name.encode(), C.byref(avg), C.byref(cnt), C.byref(tot)))
        return avg.value, cnt.value, tot.value

    def reset_stats(self):
        self.check(_lib.vt_ctx_reset_stats(self.h))

    # ---- graphs ----
    def upload(self, g: CSR) -> "DeviceGraph":
        h = C.c_void_p()
        edges = np.ascontiguousarray(g.edges)
        self.check(_lib.vt_graph_upload(self.h, g.nv, g.ne, _p(np.ascontiguousarray(g.offs, np.int64)),
                                        edges.ctypes.data_as(C.c_void_p), C.byref(h)))
        return DeviceGraph(self, h)

    def gen_rgg(self, seed: int, n: int, extra_pct: float = 0.0) -> "DeviceGraph":
        h = C.c_void_p()
        self.check(_lib.vt_gen_rgg(self.h, seed, n, extra_pct, C.byref(h)))
        return DeviceGraph(self, h)

    def gen_ba(self, seed: int, n: int, m: int) -> "DeviceGraph":
        h = C.c_void_p()
        self.check(_lib.vt_gen_ba(self.h, seed, n, m, C.byref(h)))
        return DeviceGraph(self, h)

    def gen_vite_rgg(self, n: int, x0: int = VITE_LCG_X0) -> "DeviceGraph":
        """Vite's own in-memory RGG at one rank (generateInMemGraph)."""
        h = C.c_void_p()
        self.check(_lib.vt_gen_vite_rgg(self.h, n, x0, C.byref(h)))
        return DeviceGraph(self, h)

    def compare_communities(self, truth, found) -> dict:
        """Pair counts, precision, recall, F-score and Gini coefficients of a partition
        against a ground truth (vt_compare_communities)."""
        t = np.ascontiguousarray(truth, np.int64)
        f = np.ascontiguousarray(found, np.int64)
        if len(t) != len(f):
            raise ValueError("truth and found differ in length")
        st = _CompareStats()
        self.check(_lib.vt_compare_communities(self.h, len(t), _p(t), _p(f), C.byref(st)))
        return {k: getattr(st, k) for k, _ in _CompareStats._fields_}

    def modularity(self, cw, cdeg, constant: float, sequential: bool = False) -> float:
        cw = np.ascontiguousarray(cw, np.float64)
        cdeg = np.ascontiguousarray(cdeg, np.float64)
        q = C.c_double()
        self.check(_lib.vt_modularity(self.h, len(cw), _p(cw, _f64p), _p(cdeg, _f64p), constant,
                                      VT_FLAG_SEQUENTIAL_SUMS if sequential else 0, C.byref(q)))
        return q.value


class DeviceGraph:
    def __init__(self, ctx: Context, h):
        self.ctx = ctx
        self.h = h
        nv, ne = C.c_int64(), C.c_int64()
        _lib.vt_graph_info(h, C.byref(nv), C.byref(ne))
        self.nv, self.ne = nv.value, ne.value

    def free(self):
        if self.h and self.ctx.h:
            self.ctx.check(_lib.vt_graph_free(self.ctx.h, self.h))
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def offsets(self) -> np.ndarray:
        """The CSR row offsets only (nv + 1 entries)."""
        offs = np.zeros(self.nv + 1, np.int64)
        self.ctx.check(_lib.vt_graph_download(self.ctx.h, self.h, _p(offs), None))
        return offs

    def download(self) -> CSR:
        offs = np.zeros(self.nv + 1, np.int64)
        edges = np.zeros(self.ne, EDGE_DTYPE)
        self.ctx.check(_lib.vt_graph_download(self.ctx.h, self.h, _p(offs), edges.ctypes.data_as(C.c_void_p)))
        return CSR(offs, edges)

    def init_degrees(self, sequential: bool = False):
        vdeg = np.zeros(self.nv)
        c = C.c_double()
        self.ctx.check(_lib.vt_init_degrees(self.ctx.h, self.h, VT_FLAG_SEQUENTIAL_SUMS if sequential else 0,
                                            _p(vdeg, _f64p), C.byref(c)))
        return vdeg, c.value

    def louvain_phase(self, lower=-1.0, thresh=1e-6, et_type=0, et_delta=1.0, mode="plain", num_color=0,
                      colors=None, sequential=False, max_iters=0, want_cvect=True):
        m = {"plain": VT_MODE_PLAIN, "color": VT_MODE_COLOR, "vorder": VT_MODE_VORDER}[mode]
        p = _PhaseParams(lower, thresh, et_type, m, et_delta, num_color, None,
                         VT_FLAG_SEQUENTIAL_SUMS if sequential else 0, max_iters)
        if colors is not None:
            colors = np.ascontiguousarray(colors, np.int64)
            p.colors = _p(colors)
        cvect = np.zeros(self.nv, np.int64) if want_cvect else None
        it, q = C.c_int32(), C.c_double()
        self.ctx.check(_lib.vt_louvain_phase(self.ctx.h, self.h, C.byref(p), _p(cvect) if want_cvect else None,
                                             C.byref(it), C.byref(q)))
        return q.value, cvect, it.value

    def phase_begin(self, lower=-1.0, thresh=1e-6, et_type=0, et_delta=1.0, mode="plain", num_color=0, colors=None,
                    sequential=False, max_iters=0) -> "PhaseState":
        m = {"plain": VT_MODE_PLAIN, "color": VT_MODE_COLOR, "vorder": VT_MODE_VORDER}[mode]
        p = _PhaseParams(lower, thresh, et_type, m, et_delta, num_color, None,
                         VT_FLAG_SEQUENTIAL_SUMS if sequential else 0, max_iters)
        if colors is not None:
            colors = np.ascontiguousarray(colors, np.int64)
            p.colors = _p(colors)
        h = C.c_void_p()
        self.ctx.check(_lib.vt_phase_begin(self.ctx.h, self.h, C.byref(p), C.byref(h)))
        return PhaseState(self, h)

    def last_trace(self, cap: int = 10000):
        q = np.zeros(cap)
        mv = np.zeros(cap, np.int64)
        n = C.c_int32()
        self.ctx.check(_lib.vt_last_phase_trace(self.ctx.h, cap, _p(q, _f64p), _p(mv), C.byref(n)))
        k = min(n.value, cap)
        return q[:k], mv[:k]

    def sweep(self, curr, csize, cdeg, vdeg, constant):
        nv = self.nv
        curr = np.ascontiguousarray(curr, np.int64)
        csize = np.ascontiguousarray(csize, np.int64)
        cdeg = np.ascontiguousarray(cdeg, np.float64)
        vdeg = np.ascontiguousarray(vdeg, np.float64)
        target = np.zeros(nv, np.int64)
        cw = np.zeros(nv)
        dsize = np.zeros(nv, np.int64)
        ddeg = np.zeros(nv)
        self.ctx.check(_lib.vt_sweep(self.ctx.h, self.h, _p(curr), _p(csize), _p(cdeg, _f64p), _p(vdeg, _f64p),
                                     constant, _p(target), _p(cw, _f64p), _p(dsize), _p(ddeg, _f64p)))
        return target, cw, dsize, ddeg

    def rebuild(self, cvect=None) -> "DeviceGraph":
        h = C.c_void_p()
        cp = None
        if cvect is not None:
            cvect = np.ascontiguousarray(cvect, np.int64)
            cp = _p(cvect)
        self.ctx.check(_lib.vt_rebuild(self.ctx.h, self.h, cp, C.byref(h)))
        return DeviceGraph(self.ctx, h)

    def color(self, nhash: int, target_pct: int = 70, single: bool = False):
        colors = np.zeros(self.nv, np.int64)
        nc = C.c_int64()
        self.ctx.check(_lib.vt_color(self.ctx.h, self.h, nhash, target_pct, int(single), _p(colors), C.byref(nc)))
        return nc.value, colors

    def run(self, threshold=1e-6, threshold_cycling=False, et_type=0, et_delta=1.0, coloring=False,
            vertex_order=False, max_colors=8, single_color_iter=False, one_phase=False, sequential=False):
        rp = _run_params(threshold, threshold_cycling, et_type, et_delta, coloring, vertex_order, max_colors,
                         single_color_iter, one_phase, sequential)
        res = _RunResult()
        self.ctx.check(_lib.vt_run(self.ctx.h, self.h, C.byref(rp), C.byref(res)))
        return _run_result(res)


def _run_params(threshold=1e-6, threshold_cycling=False, et_type=0, et_delta=1.0, coloring=False,
                vertex_order=False, max_colors=8, single_color_iter=False, one_phase=False, sequential=False,
                flags=0):
    return _RunParams(threshold, int(threshold_cycling), et_type, et_delta, int(coloring), int(vertex_order),
                      max_colors, int(single_color_iter), int(one_phase),
                      (VT_FLAG_SEQUENTIAL_SUMS if sequential else 0) | flags)


def _run_result(res) -> "RunResult":
    n, m = res.nlevels, res.ncalls
    return RunResult(list(res.level_q[:n]), list(res.level_iters[:n]), list(res.level_nv[:n]),
                     list(res.level_ne[:n]), list(res.call_q[:m]), list(res.call_iters[:m]), res.final_q,
                     res.total_iters, res.num_colors, res.clustering_seconds, res.total_seconds)


class Batch:
    """Many device graphs run through the phase loop together (vt_batch_*)."""

    def __init__(self, ctx: "Context", graphs):
        self.ctx = ctx
        self.n = len(graphs)
        arr = (C.c_void_p * self.n)(*[g.h for g in graphs])
        h = C.c_void_p()
        ctx.check(_lib.vt_batch_create(ctx.h, self.n, arr, C.byref(h)))
        self.h = h
        n, nv, ne = C.c_int32(), C.c_int64(), C.c_int64()
        ctx.check(_lib.vt_batch_info(self.h, C.byref(n), C.byref(nv), C.byref(ne)))
        self.total_nv, self.total_ne = nv.value, ne.value
        # the per-graph results (about 10 KB each), allocated once: vt_batch_run clears and
        # fills every entry on each run
        self._res = (_RunResult * self.n)()

    def run(self, use_graph=True, **kw):
        """Returns (list of RunResult in graph order, stats dict)."""
        rp = _run_params(flags=0 if use_graph else VT_FLAG_NO_GRAPH, **kw)
        res = self._res
        st = _BatchStats()
        self.ctx.check(_lib.vt_batch_run(self.ctx.h, self.h, C.byref(rp), res, C.byref(st)))
        return [_run_result(r) for r in res], dict(phases=st.phases, rounds=st.rounds, seconds=st.seconds,
                                                   relists=st.relists)

    def autotune(self, prep: int = 1, reps: int = 2) -> dict:
        """Picks the launch tuning for this batch's degree mix and keeps it on the context."""
        t = _Tuning()
        self.ctx.check(_lib.vt_batch_autotune(self.ctx.h, self.h, prep, reps, C.byref(t)))
        return _tuning_dict(t)

    def free(self):
        if self.h:
            h, self.h = self.h, None
            self.ctx.check(_lib.vt_batch_free(self.ctx.h, h))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


REQ_RGG, REQ_BA = 0, 1


def _request(kind, seed, n, param, **run_kw):
    k = {"rgg": REQ_RGG, "ba": REQ_BA}[kind] if isinstance(kind, str) else kind
    return _Request(k, seed, n, float(param), _run_params(**run_kw))


def request_reservation(kind, seed, n, param, **run_kw) -> int:
    return int(_lib.vt_request_reservation(C.byref(_request(kind, seed, n, param, **run_kw))))


class Server:
    """Serving loop (vt_server_*): requests admitted in arrival order by slot and HBM
    budget, run concurrently on per-slot streams."""

    def __init__(self, device: int = 0, hbm_budget: int = 200 << 30, slots: int = 4, max_batch: int = None,
                 batch_max_n: int = None):
        h = C.c_void_p()
        rc = _lib.vt_server_create(device, hbm_budget, slots, C.byref(h))
        if rc != VT_OK:
            raise VtError(rc, "vt_server_create failed")
        self.h = h
        if max_batch is not None or batch_max_n is not None:
            self.set_batching(32 if max_batch is None else max_batch, 65536 if batch_max_n is None else batch_max_n)

    def set_batching(self, max_batch: int, max_n: int):
        """Small requests queued together run as one batched job (max_batch 1: never)."""
        self._check(_lib.vt_server_set_batching(self.h, max_batch, max_n))

    def _check(self, rc):
        if rc != VT_OK:
            msg = _lib.vt_server_last_error(self.h)
            raise VtError(rc, msg.decode() if msg else "")

    def submit(self, kind, seed, n, param, **run_kw) -> int:
        t = C.c_int64()
        self._check(_lib.vt_server_submit(self.h, C.byref(_request(kind, seed, n, param, **run_kw)), C.byref(t)))
        return t.value

    def wait(self, ticket: int):
        """(RunResult, info dict) of a finished request."""
        res, info = _RunResult(), _RequestInfo()
        self._check(_lib.vt_server_wait(self.h, ticket, C.byref(res), C.byref(info)))
        return _run_result(res), {f: getattr(info, f) for f, _ in _RequestInfo._fields_}

    def close(self):
        if self.h:
            h, self.h = self.h, None
            _lib.vt_server_destroy(h)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PhaseState:
    """A Louvain phase advanced one iteration at a time, state resident on the device."""

    def __init__(self, g: DeviceGraph, h):
        self.g = g
        self.ctx = g.ctx
        self.h = h

    def step(self):
        q, done = C.c_double(), C.c_int32()
        self.ctx.check(_lib.vt_phase_step(self.ctx.h, self.h, C.byref(q), C.byref(done)))
        return q.value, bool(done.value)

    def sweep(self):
        self.ctx.check(_lib.vt_phase_sweep(self.ctx.h, self.h))

    def state(self):
        nv = self.g.nv
        curr, csize, target = np.zeros(nv, np.int64), np.zeros(nv, np.int64), np.zeros(nv, np.int64)
        cdeg, vdeg, cw = np.zeros(nv), np.zeros(nv), np.zeros(nv)
        c = C.c_double()
        self.ctx.check(_lib.vt_phase_get_state(self.ctx.h, self.h, _p(curr), _p(csize), _p(cdeg, _f64p),
                                               _p(vdeg, _f64p), C.byref(c), _p(target), _p(cw, _f64p)))
        return dict(curr=curr, csize=csize, cdeg=cdeg, vdeg=vdeg, constant=c.value, target=target, cw=cw)

    def end(self, want_cvect=True):
        cvect = np.zeros(self.g.nv, np.int64) if want_cvect else None
        it, q = C.c_int32(), C.c_double()
        h, self.h = self.h, None
        self.ctx.check(_lib.vt_phase_end(self.ctx.h, h, _p(cvect) if want_cvect else None, C.byref(it), C.byref(q)))
        return q.value, cvect, it.value

    def __del__(self):
        try:
            if self.h:
                self.end(want_cvect=False)
        except Exception:
            pass


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
    clustering_seconds: float = 0.0
    total_seconds: float = 0.0
    extra: dict = field(default_factory=dict)


def read_bin(path: str) -> CSR:
    """Vite binary graph file (distgraph.cpp:149-150,181,220)."""
    with open(path, "rb") as f:
        nv, ne = np.fromfile(f, dtype=np.int64, count=2)
        offs = np.fromfile(f, dtype=np.int64, count=int(nv) + 1)
        edges = np.fromfile(f, dtype=EDGE_DTYPE, count=int(ne))
    return CSR(offs - offs[0], edges)


def write_bin(path: str, g: CSR) -> None:
    """Vite binary graph file (writer: distgraph.cpp:1007-1045)."""
    with open(path, "wb") as f:
        np.array([g.nv, g.ne], dtype=np.int64).tofile(f)
        g.offs.astype(np.int64).tofile(f)
        np.ascontiguousarray(g.edges).tofile(f)
