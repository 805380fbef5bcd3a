"""ctypes view of libeg, the MI355X-native EverythingGraph data-parallel path.

The product is the C ABI declared in include/eg.h and implemented by
everythinggraph_amd/lib/libeg.so (HIP, gfx950) plus the C drop-in drivers in
everythinggraph_amd/bin.  This module is the thin Python handle the tests and
the benchmark use to call that ABI; it holds no compute of its own and raises
when the shared library is missing or the device is unavailable.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libeg.so")
# EG_BIN_DIR selects another build of the drop-in drivers (the sanitizer build)
BIN_DIR = os.environ.get("EG_BIN_DIR") or os.path.join(_HERE, "bin")

EG_OK = 0
EG_EINVAL = -1
EG_ENOMEM = -2
EG_EIO = -3
EG_EDEVICE = -4
EG_ERANGE = -5
EG_ENOTSUP = -6
EG_ESTATE = -7
EG_EAGAIN = -8
EG_ETIMEDOUT = -9

UNDIRECTED = 1
SYMMETRIC = 2
WEIGHTED = 4

MODE_GRID_SORT = 0
MODE_GRID_NOSORT = 1
MODE_ADJ_SORT = 2
MODE_ADJ_SORT_FULL = 3
MODE_ADJ_COUNT = 4
MODE_ADJ_DYNAMIC = 5
MODE_GRID_SORT_SRC = 6
MODE_EDGES_SRC = 7
MODE_EDGES_SRC_DST = 8
MODE_EDGES_RAW = 9

LAYOUT = {"auto": 0, "adj": 1, "grid": 2, "edges": 3}

# Records that eg_csr_build_noalloc accepts (its fixed scratch array).
NOALLOC_MAX_ENTRIES = 2097152
DANGLING = {"drop": 0, "uniform": 1}

# Exported symbols of libeg (the declarations of include/eg.h).
ABI_SYMBOLS = (
    "eg_last_error", "eg_device_query", "eg_version", "eg_graph_create",
    "eg_graph_load_file", "eg_graph_load_records", "eg_graph_load_edges",
    "eg_graph_generate", "eg_generate_edges", "eg_graph_free", "eg_graph_get_info",
    "eg_graph_get_degrees", "eg_graph_get_csr", "eg_graph_get_csc", "eg_graph_get_grid",
    "eg_graph_get_edgelist", "eg_pagerank", "eg_spmv", "eg_bfs", "eg_wcc", "eg_sssp",
    "eg_bfs_batch", "eg_ppr_batch", "eg_prof_enable", "eg_prof_reset", "eg_prof_read",
    "eg_sync", "eg_ppr", "eg_autotune", "eg_tune_set", "eg_tune_get", "eg_pr_binned_geometry",
    "eg_serve_start", "eg_serve_submit_bfs", "eg_serve_submit_ppr",
    "eg_serve_wait", "eg_serve_stats_get", "eg_serve_stop", "eg_serve_replay",
    "eg_csr_build_noalloc", "eg_bfs_batch_info_get", "eg_graph_prep_info",
)


class EgError(RuntimeError):
    def __init__(self, code: int, what: str, msg: str):
        super().__init__(f"{what} failed ({code}): {msg}")
        self.code = code


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("compute_units", ctypes.c_int32),
                ("gcn_arch", ctypes.c_int32), ("hbm_bytes", ctypes.c_uint64),
                ("free_bytes", ctypes.c_uint64), ("peak_hbm_gbs", ctypes.c_double)]


class GraphInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("P", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("load_mode", ctypes.c_int32), ("m_in", ctypes.c_uint64), ("m", ctypes.c_uint64),
                ("has_csr", ctypes.c_int32), ("has_csc", ctypes.c_int32),
                ("has_grid", ctypes.c_int32), ("has_edgelist", ctypes.c_int32),
                ("build_ms", ctypes.c_double), ("device_bytes", ctypes.c_uint64)]


class PrepInfo(ctypes.Structure):
    _fields_ = [("layout_ms", ctypes.c_double), ("order_ms", ctypes.c_double),
                ("binned_plan_ms", ctypes.c_double), ("bfs_rows_ms", ctypes.c_double),
                ("layout_bytes", ctypes.c_uint64), ("order_bytes", ctypes.c_uint64),
                ("binned_plan_bytes", ctypes.c_uint64), ("bfs_rows_bytes", ctypes.c_uint64)]


class GenParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("scale", ctypes.c_int32), ("m_in", ctypes.c_uint64),
                ("seed", ctypes.c_uint64), ("a", ctypes.c_double), ("b", ctypes.c_double),
                ("c", ctypes.c_double), ("weight_seed", ctypes.c_uint64)]


class PrOpts(ctypes.Structure):
    _fields_ = [("damping", ctypes.c_double), ("iterations", ctypes.c_int32),
                ("dangling", ctypes.c_int32), ("init", ctypes.c_double), ("layout", ctypes.c_int32)]


class BfsStats(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int32), ("pull_levels", ctypes.c_int32),
                ("reached", ctypes.c_uint64), ("edges_traversed", ctypes.c_uint64),
                ("device_ms", ctypes.c_double)]


class BfsBatchInfo(ctypes.Structure):
    _fields_ = [("traversals", ctypes.c_uint32), ("lanes", ctypes.c_uint32),
                ("levels", ctypes.c_uint32), ("push_levels", ctypes.c_uint32),
                ("host_syncs", ctypes.c_uint32), ("fallbacks", ctypes.c_uint32),
                ("device_ms", ctypes.c_double)]


class SsspStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("pull_iterations", ctypes.c_int32),
                ("reached", ctypes.c_uint64), ("device_ms", ctypes.c_double)]


SSSP_WEIGHTS = {"ref": 0, "unit": 1}
SSSP_UNREACHED = 0xFFFFFFFE


class KernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_uint64),
                ("total_ms", ctypes.c_double)]


REQ_BFS, REQ_PPR = 0, 1


class ServeOpts(ctypes.Structure):
    _fields_ = [("bfs_lanes", ctypes.c_uint32), ("ppr_columns", ctypes.c_uint32),
                ("hbm_reserve", ctypes.c_uint64), ("ppr_alpha", ctypes.c_double),
                ("ppr_tol", ctypes.c_double), ("ppr_max_iter", ctypes.c_int32),
                ("plain", ctypes.c_int32), ("max_pending", ctypes.c_uint32),
                ("ppr_grid", ctypes.c_uint32), ("idle_plain", ctypes.c_int32),
                ("pad", ctypes.c_uint32)]


class ServeResult(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("status", ctypes.c_int32),
                ("levels", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("reached", ctypes.c_uint64), ("delta", ctypes.c_double),
                ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double)]


class ServeStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("failed", ctypes.c_uint64), ("pending", ctypes.c_uint32),
                ("bfs_inflight", ctypes.c_uint32), ("ppr_inflight", ctypes.c_uint32),
                ("bfs_lanes", ctypes.c_uint32), ("ppr_columns", ctypes.c_uint32),
                ("pad", ctypes.c_uint32), ("bfs_levels", ctypes.c_uint64),
                ("ppr_steps", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64),
                ("served_plain", ctypes.c_uint64)]


class TuneEntry(ctypes.Structure):
    _fields_ = [("param", ctypes.c_char * 24), ("value", ctypes.c_int32),
                ("chosen", ctypes.c_int32), ("ms", ctypes.c_double)]


TUNE_PULL, TUNE_SPMM, TUNE_MSBFS, TUNE_GRID = 1, 2, 4, 8


class ReplayParams(ctypes.Structure):
    _fields_ = [("qps", ctypes.c_double), ("requests", ctypes.c_uint32),
                ("ppr_max_seeds", ctypes.c_uint32), ("bfs_fraction", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


class ReplayReport(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint32), ("failed", ctypes.c_uint32),
                ("bfs", ctypes.c_uint32), ("ppr", ctypes.c_uint32),
                ("wall_s", ctypes.c_double), ("qps", ctypes.c_double),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("bfs_p50_ms", ctypes.c_double), ("bfs_p99_ms", ctypes.c_double),
                ("ppr_p50_ms", ctypes.c_double), ("ppr_p99_ms", ctypes.c_double),
                ("stats", ServeStats)]


def _struct_dict(st) -> dict:
    out = {}
    for name, _ in st._fields_:
        if name == "pad":
            continue
        v = getattr(st, name)
        out[name] = _struct_dict(v) if isinstance(v, ctypes.Structure) else v
    return out


def serve_opts(bfs_lanes=0, ppr_columns=0, hbm_reserve=0, ppr_alpha=0.0, ppr_tol=0.0,
               ppr_max_iter=0, plain=False, max_pending=0, ppr_grid=0,
               idle_plain=0) -> ServeOpts:
    """Serving options; zeros take the library defaults (64 lanes, 64 columns,
    1 GiB reserve, alpha 0.15, tolerance 1e-6, 200 iterations, an SpMM grid of
    8 workgroups per compute unit, idle engine serves up to 2 waiting queries
    through the plain path; idle_plain=-1 always batches)."""
    return ServeOpts(bfs_lanes, ppr_columns, hbm_reserve, ppr_alpha, ppr_tol, ppr_max_iter,
                     int(plain), max_pending, ppr_grid, idle_plain, 0)


_lib: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    """Loads libeg.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        L.eg_last_error.restype = ctypes.c_char_p
        L.eg_version.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(rc: int, what: str):
    if rc != EG_OK:
        raise EgError(rc, what, lib().eg_last_error().decode(errors="replace"))


def device_info() -> DeviceInfo:
    d = DeviceInfo()
    _check(lib().eg_device_query(ctypes.byref(d)), "eg_device_query")
    return d


def rmat_params(scale: int, edge_factor: int = 16, seed: int = 1, a=0.57, b=0.19, c=0.19,
                weight_seed: int = 1) -> GenParams:
    return GenParams(0, scale, (1 << scale) * edge_factor, seed, a, b, c, weight_seed)


def er_params(m_in: int, seed: int = 1, weight_seed: int = 1) -> GenParams:
    return GenParams(1, 0, m_in, seed, 0.0, 0.0, 0.0, weight_seed)


def generate_edges(p: GenParams, n: int, weighted: bool = False):
    """Device generator, records copied to host arrays."""
    src = np.empty(p.m_in, np.uint32)
    dst = np.empty(p.m_in, np.uint32)
    w = np.empty(p.m_in, np.float32) if weighted else None
    _check(lib().eg_generate_edges(ctypes.byref(p), ctypes.c_uint32(n), _ptr(src), _ptr(dst),
                                   _ptr(w)), "eg_generate_edges")
    return src, dst, w


class Graph:
    """A device-resident graph in one of the reference's load modes."""

    def __init__(self, n: int, load_mode: int = MODE_ADJ_SORT, undirected: bool = False,
                 symmetric: bool = False, weighted: bool = False, P: int = 0):
        self.n = int(n)
        self.flags = (UNDIRECTED if undirected else 0) | (SYMMETRIC if symmetric else 0) | (
            WEIGHTED if weighted else 0)
        self.weighted = weighted
        h = ctypes.c_void_p()
        _check(lib().eg_graph_create(ctypes.c_uint32(self.n), ctypes.c_uint32(self.flags),
                                     ctypes.c_int32(load_mode), ctypes.c_uint32(P),
                                     ctypes.byref(h)), "eg_graph_create")
        self._h = h

    # -- lifecycle --
    def close(self):
        if self._h:
            lib().eg_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def load_file(self, path: str):
        _check(lib().eg_graph_load_file(self._h, os.fsencode(path)), "eg_graph_load_file")
        return self

    def load_records(self, data: bytes):
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        _check(lib().eg_graph_load_records(self._h, buf, ctypes.c_uint64(len(data))),
               "eg_graph_load_records")
        return self

    def load_edges(self, src: np.ndarray, dst: np.ndarray, w: Optional[np.ndarray] = None):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        w = None if w is None else np.ascontiguousarray(w, np.float32)
        _check(lib().eg_graph_load_edges(self._h, ctypes.c_uint64(src.size), _ptr(src), _ptr(dst),
                                         _ptr(w)), "eg_graph_load_edges")
        return self

    def generate(self, p: GenParams):
        _check(lib().eg_graph_generate(self._h, ctypes.byref(p)), "eg_graph_generate")
        return self

    @property
    def info(self) -> GraphInfo:
        i = GraphInfo()
        _check(lib().eg_graph_get_info(self._h, ctypes.byref(i)), "eg_graph_get_info")
        return i

    @property
    def prep(self) -> dict:
        """One-time preparations built so far: ms and device bytes of the
        layout, the degree-ordered copy, the binned plan, the BFS rows."""
        i = PrepInfo()
        _check(lib().eg_graph_prep_info(self._h, ctypes.byref(i)), "eg_graph_prep_info")
        return _struct_dict(i)

    # -- layouts --
    def degrees(self):
        o = np.empty(self.n, np.uint32)
        i = np.empty(self.n, np.uint32)
        _check(lib().eg_graph_get_degrees(self._h, _ptr(o), _ptr(i)), "eg_graph_get_degrees")
        return o, i

    def csr(self):
        m = self.info.m
        ptr = np.empty(self.n + 1, np.uint64)
        col = np.empty(m, np.uint32)
        w = np.empty(m, np.float32) if self.weighted else None
        _check(lib().eg_graph_get_csr(self._h, _ptr(ptr), _ptr(col), _ptr(w)), "eg_graph_get_csr")
        return ptr, col, w

    def csc(self):
        m = self.info.m
        ptr = np.empty(self.n + 1, np.uint64)
        row = np.empty(m, np.uint32)
        w = np.empty(m, np.float32) if self.weighted else None
        _check(lib().eg_graph_get_csc(self._h, _ptr(ptr), _ptr(row), _ptr(w)), "eg_graph_get_csc")
        return ptr, row, w

    def grid(self):
        inf = self.info
        cp = np.empty(inf.P * inf.P + 1, np.uint64)
        s = np.empty(inf.m, np.uint32)
        d = np.empty(inf.m, np.uint32)
        w = np.empty(inf.m, np.float32) if self.weighted else None
        _check(lib().eg_graph_get_grid(self._h, _ptr(cp), _ptr(s), _ptr(d), _ptr(w)),
               "eg_graph_get_grid")
        return cp, s, d, w

    def edgelist(self):
        m = self.info.m
        s = np.empty(m, np.uint32)
        d = np.empty(m, np.uint32)
        w = np.empty(m, np.float32) if self.weighted else None
        _check(lib().eg_graph_get_edgelist(self._h, _ptr(s), _ptr(d), _ptr(w)),
               "eg_graph_get_edgelist")
        return s, d, w

    # -- algorithms --
    def pagerank(self, damping=0.85, iterations=10, dangling="drop", init=0.0, layout="auto",
                 download=True):
        o = PrOpts(damping, iterations, DANGLING[dangling], init, LAYOUT[layout])
        r = np.empty(self.n, np.float32) if download else None
        dh = np.zeros(max(iterations, 1), np.float64)
        lh = np.zeros(max(iterations, 1), np.float64)
        _check(lib().eg_pagerank(self._h, ctypes.byref(o), _ptr(r), _ptr(dh), _ptr(lh)),
               "eg_pagerank")
        return r, dh[:iterations], lh[:iterations]

    def spmv(self, x: np.ndarray, layout="auto"):
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty(self.n, np.float32)
        _check(lib().eg_spmv(self._h, ctypes.c_int32(LAYOUT[layout]), _ptr(x), _ptr(y)), "eg_spmv")
        return y

    def bfs(self, root: int, start_pull=False, switch=True, parents=True, download=True):
        dist = np.empty(self.n, np.uint32) if download else None
        par = np.empty(self.n, np.uint32) if (parents and download) else None
        st = BfsStats()
        _check(lib().eg_bfs(self._h, ctypes.c_uint32(root), ctypes.c_int32(int(start_pull)),
                            ctypes.c_int32(int(switch)), _ptr(dist), _ptr(par), ctypes.byref(st)),
               "eg_bfs")
        return dist, par, st

    def sssp(self, root: int, start_pull=False, switch=True, weights="ref", parents=True,
             download=True):
        """(dist, parent or None, SsspStats); weights "ref" = (u%10)+(v%10), "unit" = 1."""
        dist = np.empty(self.n, np.uint32) if download else None
        par = np.empty(self.n, np.uint32) if (parents and download) else None
        st = SsspStats()
        _check(lib().eg_sssp(self._h, ctypes.c_uint32(root), ctypes.c_int32(int(start_pull)),
                             ctypes.c_int32(int(switch)), ctypes.c_int32(SSSP_WEIGHTS[weights]),
                             _ptr(dist), _ptr(par), ctypes.byref(st)), "eg_sssp")
        return dist, par, st

    def wcc(self):
        lab = np.empty(self.n, np.int64)
        rounds = ctypes.c_int32()
        _check(lib().eg_wcc(self._h, _ptr(lab), ctypes.byref(rounds)), "eg_wcc")
        return lab, rounds.value

    def bfs_batch(self, roots: Sequence[int], download=True):
        r = np.ascontiguousarray(roots, np.uint32)
        out = np.empty((r.size, self.n), np.uint32) if download else None
        ms = ctypes.c_double()
        _check(lib().eg_bfs_batch(self._h, ctypes.c_uint32(r.size), _ptr(r), _ptr(out),
                                  ctypes.byref(ms)), "eg_bfs_batch")
        return out, ms.value

    def bfs_batch_info(self) -> dict:
        """Figures of the last bfs_batch call: traversals, lanes, levels,
        push_levels, host_syncs, fallbacks, device_ms."""
        i = BfsBatchInfo()
        _check(lib().eg_bfs_batch_info_get(self._h, ctypes.byref(i)), "eg_bfs_batch_info_get")
        return _struct_dict(i)

    def ppr_batch(self, seeds: Sequence[Sequence[int]], alpha=0.15, iterations=30, download=True):
        ptr = np.zeros(len(seeds) + 1, np.uint32)
        ptr[1:] = np.cumsum([len(s) for s in seeds])
        flat = np.ascontiguousarray(np.concatenate([np.asarray(s, np.uint32) for s in seeds])
                                    if seeds else np.zeros(0, np.uint32), np.uint32)
        out = np.empty((len(seeds), self.n), np.float32) if download else None
        ms = ctypes.c_double()
        _check(lib().eg_ppr_batch(self._h, ctypes.c_uint32(len(seeds)), _ptr(ptr), _ptr(flat),
                                  ctypes.c_double(alpha), ctypes.c_int32(iterations), _ptr(out),
                                  ctypes.byref(ms)), "eg_ppr_batch")
        return out, ms.value

    def ppr(self, seeds: Sequence[int], alpha=0.15, tol=1e-6, max_iter=200, download=True):
        """Single-vector personalized PageRank: (ranks or None, iterations, last L1 change)."""
        sd = np.ascontiguousarray(seeds, np.uint32)
        out = np.empty(self.n, np.float32) if download else None
        it, dl = ctypes.c_int32(), ctypes.c_double()
        _check(lib().eg_ppr(self._h, _ptr(sd), ctypes.c_uint32(sd.size), ctypes.c_double(alpha),
                            ctypes.c_double(tol), ctypes.c_int32(max_iter), _ptr(out),
                            ctypes.byref(it), ctypes.byref(dl)), "eg_ppr")
        return out, it.value, dl.value

    # -- kernel shapes --
    def autotune(self, kernels=TUNE_PULL | TUNE_SPMM | TUNE_MSBFS) -> list:
        """Times every variant of the selected kernels on this graph and keeps
        the fastest: [(param, value, ms, chosen), ...]."""
        arr = (TuneEntry * 32)()
        cnt = ctypes.c_int32()
        _check(lib().eg_autotune(self._h, ctypes.c_uint32(kernels), arr, ctypes.c_int32(32),
                                 ctypes.byref(cnt)), "eg_autotune")
        return [(arr[i].param.decode(), arr[i].value, arr[i].ms, bool(arr[i].chosen))
                for i in range(min(cnt.value, 32))]

    def binned_geometry(self) -> dict:
        """Geometry of the binned PageRank plan (after a binned pagerank call)."""
        out = np.zeros(7, np.uint32)
        _check(lib().eg_pr_binned_geometry(self._h, _ptr(out)), "eg_pr_binned_geometry")
        return dict(zip(("bins", "chunks", "tiles", "items", "long_rows", "segments", "hot"),
                        (int(v) for v in out)))

    def tune_set(self, param: str, value: int):
        _check(lib().eg_tune_set(self._h, param.encode(), ctypes.c_int32(value)), "eg_tune_set")

    def tune_get(self, param: str) -> int:
        v = ctypes.c_int32()
        _check(lib().eg_tune_get(self._h, param.encode(), ctypes.byref(v)), "eg_tune_get")
        return v.value

    # -- serving --
    def serve(self, **opts) -> "Server":
        return Server(self, serve_opts(**opts))

    def serve_replay(self, qps: float, requests: int, bfs_fraction=0.8, ppr_max_seeds=8,
                     seed=4, **opts) -> dict:
        r = ReplayReport()
        p = ReplayParams(qps, requests, ppr_max_seeds, bfs_fraction, seed)
        _check(lib().eg_serve_replay(self._h, ctypes.byref(serve_opts(**opts)), ctypes.byref(p),
                                     ctypes.byref(r)), "eg_serve_replay")
        return _struct_dict(r)

    # -- timing --
    def prof_enable(self, on=True):
        _check(lib().eg_prof_enable(self._h, ctypes.c_int32(int(on))), "eg_prof_enable")

    def prof_reset(self):
        _check(lib().eg_prof_reset(self._h), "eg_prof_reset")

    def prof_read(self) -> dict:
        arr = (KernelTime * 64)()
        cnt = ctypes.c_int32()
        _check(lib().eg_prof_read(self._h, arr, ctypes.c_int32(64), ctypes.byref(cnt)),
               "eg_prof_read")
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms)
                for i in range(min(cnt.value, 64))}

    def sync(self):
        _check(lib().eg_sync(self._h), "eg_sync")


class Server:
    """Continuous-batching server over a loaded Graph (eg_serve_*).  The graph
    cannot be used by other calls until stop()."""

    def __init__(self, g: Graph, opts: ServeOpts):
        self.g = g
        self._h = ctypes.c_void_p()
        self._out = {}
        _check(lib().eg_serve_start(g._h, ctypes.byref(opts), ctypes.byref(self._h)),
               "eg_serve_start")

    def submit_bfs(self, root: int, download=True) -> int:
        out = np.empty(self.g.n, np.uint32) if download else None
        t = ctypes.c_uint64()
        _check(lib().eg_serve_submit_bfs(self._h, ctypes.c_uint32(root), _ptr(out),
                                         ctypes.byref(t)), "eg_serve_submit_bfs")
        self._out[t.value] = out
        return t.value

    def submit_ppr(self, seeds: Sequence[int], download=True) -> int:
        sd = np.ascontiguousarray(seeds, np.uint32)
        out = np.empty(self.g.n, np.float32) if download else None
        t = ctypes.c_uint64()
        _check(lib().eg_serve_submit_ppr(self._h, _ptr(sd), ctypes.c_uint32(sd.size), _ptr(out),
                                         ctypes.byref(t)), "eg_serve_submit_ppr")
        self._out[t.value] = out
        return t.value

    def wait(self, ticket: int, timeout_ms: float = -1.0):
        """(result dict, output array or None); raises EgError if the query
        failed or the wait timed out."""
        r = ServeResult()
        _check(lib().eg_serve_wait(self._h, ctypes.c_uint64(ticket), ctypes.c_double(timeout_ms),
                                   ctypes.byref(r)), "eg_serve_wait")
        out = self._out.pop(ticket, None)
        if r.status != EG_OK:
            raise EgError(r.status, "query", lib().eg_last_error().decode(errors="replace"))
        return _struct_dict(r), out

    def stats(self) -> dict:
        st = ServeStats()
        _check(lib().eg_serve_stats_get(self._h, ctypes.byref(st)), "eg_serve_stats_get")
        return _struct_dict(st)

    def stop(self):
        if self._h:
            h, self._h = self._h, ctypes.c_void_p()
            _check(lib().eg_serve_stop(h), "eg_serve_stop")
        self._out.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()


def sample_roots(deg: np.ndarray, k: int, seed: int = 2) -> np.ndarray:
    """k distinct vertices with degree > 0, drawn from the counter-based stream
    mix64(mix64(seed) ^ j) mapped to [0, n) by multiply-shift (the same mixing
    function as the generator)."""
    n = deg.size
    M = (1 << 64) - 1

    def mix(z):
        z = (z + 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    key = mix(seed)
    out, seen = [], set()
    j = 0
    if int((deg > 0).sum()) < k:
        raise ValueError("not enough vertices with edges")
    while len(out) < k:
        h = mix(key ^ j)
        v = ((h >> 32) * n) >> 32
        j += 1
        if deg[v] > 0 and v not in seen:
            seen.add(v)
            out.append(v)
    return np.asarray(out, np.uint32)
