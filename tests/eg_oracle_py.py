"""numpy view of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: it is only ever the checker here."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# EG_ORACLE_SO selects another build of the same source (the sanitizer build)
ORACLE_SO = os.environ.get("EG_ORACLE_SO") or os.path.join(ORACLE_DIR, "liboracle.so")

_L = None


def L():
    global _L
    if _L is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        _L = ctypes.CDLL(ORACLE_SO)
        _L.ego_mix64.restype = ctypes.c_uint64
        _L.ego_mix64.argtypes = [ctypes.c_uint64]
        _L.ego_partition_id.restype = ctypes.c_uint32
        _L.ego_partition_id.argtypes = [ctypes.c_uint32] * 3
        _L.ego_default_P.restype = ctypes.c_uint32
        _L.ego_default_P.argtypes = [ctypes.c_uint32]
    return _L


def P(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


U32, U64, I32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32
DBL = ctypes.c_double

RMAT_ABC = (0.57, 0.19, 0.19)


def thresholds(a=0.57, b=0.19, c=0.19):
    two32 = 4294967296.0
    return int(a * two32), int((a + b) * two32), int((a + b + c) * two32)


def gen_rmat(scale, m_in, seed=1, abc=RMAT_ABC):
    s = np.empty(m_in, np.uint32)
    d = np.empty(m_in, np.uint32)
    ta, tab, tabc = thresholds(*abc)
    L().ego_gen_rmat(I32(scale), U64(m_in), U64(seed), U32(ta), U32(tab), U32(tabc), P(s), P(d))
    return s, d


def gen_er(n, m_in, seed=1):
    s = np.empty(m_in, np.uint32)
    d = np.empty(m_in, np.uint32)
    L().ego_gen_er(U32(n), U64(m_in), U64(seed), P(s), P(d))
    return s, d


def gen_weights(m_in, seed=1):
    w = np.empty(m_in, np.float32)
    L().ego_gen_weights(U64(m_in), U64(seed), P(w))
    return w


def symmetrize(src, dst, w=None):
    m = src.size
    os_ = np.empty(2 * m, np.uint32)
    od = np.empty(2 * m, np.uint32)
    ow = np.empty(2 * m, np.float32) if w is not None else None
    L().ego_symmetrize(U64(m), P(src), P(dst), P(w), P(os_), P(od), P(ow))
    return os_, od, ow


def partition_id(n, Pp, v):
    return L().ego_partition_id(n, Pp, v)


def default_P(n):
    return L().ego_default_P(n)


def build_csr(n, src, dst, w=None):
    m = src.size
    ptr = np.empty(n + 1, np.uint64)
    col = np.empty(m, np.uint32)
    ow = np.empty(m, np.float32) if w is not None else None
    L().ego_build_csr(U32(n), U64(m), P(src), P(dst), P(w), P(ptr), P(col), P(ow))
    return ptr, col, ow


def build_csc(n, ptr, col, w=None):
    m = col.size
    cp = np.empty(n + 1, np.uint64)
    row = np.empty(m, np.uint32)
    ow = np.empty(m, np.float32) if w is not None else None
    L().ego_build_csc(U32(n), U64(m), P(ptr), P(col), P(w), P(cp), P(row), P(ow))
    return cp, row, ow


def build_grid(n, Pp, src, dst, w=None, presort=True):
    m = src.size
    cp = np.empty(Pp * Pp + 1, np.uint64)
    gs = np.empty(m, np.uint32)
    gd = np.empty(m, np.uint32)
    gw = np.empty(m, np.float32) if w is not None else None
    L().ego_build_grid(U32(n), U32(Pp), U64(m), P(src), P(dst), P(w), I32(int(presort)), P(cp),
                       P(gs), P(gd), P(gw))
    return cp, gs, gd, gw


def build_edgelist(n, src, dst, w, mode):
    m = src.size
    os_ = np.empty(m, np.uint32)
    od = np.empty(m, np.uint32)
    ow = np.empty(m, np.float32) if w is not None else None
    L().ego_build_edgelist(U32(n), U64(m), P(src), P(dst), P(w), I32(mode), P(os_), P(od), P(ow))
    return os_, od, ow


def pagerank_pull(n, in_ptr, in_idx, out_deg, damping=0.85, iterations=20, init=None,
                  dangling_uniform=False):
    r = np.empty(n, np.float64)
    dh = np.zeros(max(iterations, 1), np.float64)
    lh = np.zeros(max(iterations, 1), np.float64)
    L().ego_pagerank_pull(U32(n), P(in_ptr), P(in_idx), P(out_deg), DBL(damping), I32(iterations),
                          DBL(1.0 / n if init is None else init), I32(int(dangling_uniform)), P(r),
                          P(dh), P(lh))
    return r, dh[:iterations], lh[:iterations]


def pagerank_grid(n, Pp, cp, gs, gd, out_deg, damping=0.85, iterations=20, init=None):
    r = np.empty(n, np.float64)
    L().ego_pagerank_grid(U32(n), U32(Pp), P(cp), P(gs), P(gd), P(out_deg), DBL(damping),
                          I32(iterations), DBL(1.0 / n if init is None else init), P(r))
    return r


def spmv_pull(n, in_ptr, in_idx, in_w, x):
    y = np.empty(n, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    L().ego_spmv_pull(U32(n), P(in_ptr), P(in_idx), P(in_w), P(x), P(y))
    return y


def spmv_grid(n, Pp, cp, gs, gd, gw, x):
    y = np.empty(n, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    L().ego_spmv_grid(U32(n), U32(Pp), P(cp), P(gs), P(gd), P(gw), P(x), P(y))
    return y


def bfs(n, m, out_ptr, out_col, in_ptr, in_idx, root, start_pull=False, switch=True):
    dist = np.empty(n, np.uint32)
    ls = np.zeros(4096, np.uint64)
    lv = L().ego_bfs(U32(n), U64(m), P(out_ptr), P(out_col), P(in_ptr), P(in_idx), U32(root),
                     I32(int(start_pull)), I32(int(switch)), P(dist), P(ls), I32(4096))
    return dist, lv, ls[:lv]


def bfs_grid(n, Pp, cp, gs, gd, root):
    dist = np.empty(n, np.uint32)
    lv = L().ego_bfs_grid(U32(n), U32(Pp), P(cp), P(gs), P(gd), U32(root), P(dist))
    return dist, lv


def bfs_edgelist(n, src, dst, root):
    dist = np.empty(n, np.uint32)
    lv = L().ego_bfs_edgelist(U32(n), U64(src.size), P(src), P(dst), U32(root), P(dist))
    return dist, lv


def wcc_push(n, out_ptr, out_col):
    lab = np.empty(n, np.int64)
    rounds = L().ego_wcc_push(U32(n), P(out_ptr), P(out_col), P(lab))
    return lab, rounds


def wcc_edges(n, src, dst):
    lab = np.empty(n, np.int64)
    rounds = L().ego_wcc_edges(U32(n), U64(src.size), P(src), P(dst), P(lab))
    return lab, rounds


def ppr(n, in_ptr, in_idx, out_deg, seeds, alpha=0.15, iterations=30):
    s = np.ascontiguousarray(seeds, np.uint32)
    out = np.empty(n, np.float64)
    L().ego_ppr(U32(n), P(in_ptr), P(in_idx), P(out_deg), U32(s.size), P(s), DBL(alpha),
                I32(iterations), P(out))
    return out


def check_bfs_parents(n, out_ptr, out_col, root, dist, parent):
    return L().ego_check_bfs_parents(U32(n), P(out_ptr), P(out_col), U32(root), P(dist), P(parent))


SSSP_UNREACHED = 0xFFFFFFFE  # UINT32_MAX - 1, sssp_pushpull.c:8


def sssp_weight(u, v, kind=0):
    """Edge weight of the SSSP oracle (numpy-vectorised twin of ego_sssp_weight)."""
    u = np.asarray(u, np.uint64)
    v = np.asarray(v, np.uint64)
    if kind == 1:
        return np.ones(np.broadcast(u, v).shape, np.uint64)
    return u % 10 + v % 10 + (1 if kind == 2 else 0)


def sssp(n, out_ptr, out_col, root, kind=0):
    dist = np.empty(n, np.uint32)
    rounds = L().ego_sssp(U32(n), P(out_ptr), P(out_col), U32(root), I32(kind), P(dist))
    return dist, rounds


def check_sssp_parents(n, out_ptr, out_col, root, kind, dist, parent):
    return L().ego_check_sssp_parents(U32(n), P(out_ptr), P(out_col), U32(root), I32(kind),
                                      P(dist), P(parent))


def out_degree(n, src):
    return np.bincount(src, minlength=n).astype(np.uint32)


def record_row_sums(n, src, dst, both):
    """(count, column sum, column xor) of every row, from the records."""
    cnt = np.empty(n, np.uint64)
    sm = np.empty(n, np.uint64)
    xr = np.empty(n, np.uint32)
    L().ego_record_row_sums(U32(n), U64(src.size), P(src), P(dst), I32(int(both)), P(cnt),
                            P(sm), P(xr))
    return cnt, sm, xr


def csr_row_sums(n, ptr, col):
    """(column sum, column xor) of every row of a CSR."""
    sm = np.empty(n, np.uint64)
    xr = np.empty(n, np.uint32)
    L().ego_csr_row_sums(U32(n), P(ptr), P(col), P(sm), P(xr))
    return sm, xr


def rows_of(n, src, dst, both, sel):
    """Rows of the ascending, distinct vertices sel in stored order: (ptr, col)."""
    sel = np.ascontiguousarray(sel, np.uint32)
    ptr = np.empty(sel.size + 1, np.uint64)
    L().ego_rows_of(U32(n), U64(src.size), P(src), P(dst), I32(int(both)), U32(sel.size),
                    P(sel), P(ptr), None)
    col = np.empty(int(ptr[-1]), np.uint32)
    L().ego_rows_of(U32(n), U64(src.size), P(src), P(dst), I32(int(both)), U32(sel.size),
                    P(sel), P(ptr), P(col))
    return ptr, col


def check_bfs_levels(n, src, dst, both, root, dist):
    """Violations of the single-source BFS level invariants over the records."""
    f = L().ego_check_bfs_levels
    f.restype = ctypes.c_int64
    return f(U32(n), U64(src.size), P(src), P(dst), I32(int(both)), U32(root), P(dist))
