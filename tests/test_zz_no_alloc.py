"""eg_csr_build_noalloc, the CSR build that allocates no device memory, called
through the C ABI with device buffers from the HIP runtime that libeg itself
loaded.  Its output must equal, bit for bit, the CSR that eg_graph_load_edges
builds in load mode 2 (init_adj_sort, init_all.c:629-697) from the same
records, and the oracle's; twice in a row, the second call smaller (the fixed
scratch array keeps the first call's contents), then with no records, on a
stream of the caller's, and past the size limit (refused)."""
import ctypes

import numpy as np
import pytest

import eg_oracle_py as O
import everythinggraph_amd as eg

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip_runtime():
    """The libamdhip64 that libeg.so linked (found in this process's maps)."""
    eg.lib()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64.so" in ln}
    assert len(paths) == 1, paths
    H = ctypes.CDLL(paths.pop())
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    H.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    H.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return H


class Dev:
    """A device buffer holding a copy of a host array (or len bytes)."""

    def __init__(self, H, nbytes):
        self.H = H
        self.p = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.p), max(nbytes, 4)) == 0
        self.nbytes = nbytes

    @classmethod
    def of(cls, H, a):
        d = cls(H, a.nbytes)
        if a.nbytes:
            assert H.hipMemcpy(d.p, a.ctypes.data, a.nbytes, H2D) == 0
        return d

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        if out.nbytes:
            assert self.H.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.p)


def build_noalloc(H, n, s, d, stream=None):
    """(row_ptr, col) from eg_csr_build_noalloc; status code when it refuses."""
    ds, dd = Dev.of(H, s), Dev.of(H, d)
    dp, dc = Dev(H, (n + 1) * 4), Dev(H, s.size * 4)
    print(f"no-alloc CSR build: n={n} m={s.size}", flush=True)
    rc = eg.lib().eg_csr_build_noalloc(ctypes.c_uint32(n), ctypes.c_uint64(s.size), ds.p, dd.p,
                                       dp.p, dc.p, stream)
    try:
        if rc != eg.EG_OK:
            return rc
        assert H.hipStreamSynchronize(stream) == 0
        return dp.get(np.uint32, n + 1), dc.get(np.uint32, s.size)
    finally:
        for b in (ds, dd, dp, dc):
            b.free()


def reference_csr(n, s, d):
    with eg.Graph(n, eg.MODE_ADJ_SORT, symmetric=True) as g:
        g.load_edges(s, d)
        ptr, col, _ = g.csr()
    return ptr, col


def check(n, s, d, got):
    ptr, col = got
    rptr, rcol = reference_csr(n, s, d)  # the existing entry point
    assert np.array_equal(ptr.astype(np.uint64), rptr) and np.array_equal(col, rcol)
    optr, ocol, _ = O.build_csr(n, s, d)
    assert np.array_equal(ptr.astype(np.uint64), optr) and np.array_equal(col, ocol)


def test_noalloc_csr_equals_load_edges():
    H = hip_runtime()
    # R-MAT 16 -u: 2 x 2^20 = 2^21 entries, exactly the limit
    s0, d0 = O.gen_rmat(16, 1 << 20, seed=5)
    s, d, _ = O.symmetrize(s0, d0)
    assert s.size == eg.NOALLOC_MAX_ENTRIES
    check(1 << 16, s, d, build_noalloc(H, 1 << 16, s, d))
    # a second, smaller call: ragged size, ids past 2^16 (three digit passes)
    rng = np.random.default_rng(6)
    n2, m2 = (1 << 20) + 7, 50001
    s2 = rng.integers(0, n2, m2).astype(np.uint32)
    d2 = rng.integers(0, n2, m2).astype(np.uint32)
    check(n2, s2, d2, build_noalloc(H, n2, s2, d2))


def test_noalloc_csr_stream_empty_and_limit():
    H = hip_runtime()
    st = ctypes.c_void_p()
    assert H.hipStreamCreate(ctypes.byref(st)) == 0
    try:
        s, d = O.gen_rmat(12, 16 << 12, seed=8)
        check(1 << 12, s, d, build_noalloc(H, 1 << 12, s, d, st))
        e = np.zeros(0, np.uint32)
        ptr, col = build_noalloc(H, 100, e, e, st)
        assert col.size == 0 and np.all(ptr == 0)
    finally:
        H.hipStreamDestroy(st)
    # one entry past the limit is refused before any device work
    rc = eg.lib().eg_csr_build_noalloc(ctypes.c_uint32(16), ctypes.c_uint64(eg.NOALLOC_MAX_ENTRIES + 1),
                                       None, None, None, None, None)
    assert rc == eg.EG_EINVAL
    assert b"exceed" in eg.lib().eg_last_error()
