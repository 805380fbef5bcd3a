"""The committed R-MAT 10 fixtures (tests/golden/, written by
tests/golden/make_fixtures.py).

* CPU: the files still have the SHA-256 of MANIFEST.json, and the oracle
  still computes them from the generator: a change to the generator or to
  any oracle function listed there fails here even if it changed the tests'
  other expectations along with it.
* GPU: libeg, loading the record file through the C ABI exactly as the
  drop-in drivers do, gives the fixtures' layouts, BFS levels and WCC labels
  bit for bit, and PageRank within 1e-6 L1 (and 1e-6 of the largest rank
  elementwise) of the FP64 ranks.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import eg_oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
N = 1 << 10


def manifest():
    with open(os.path.join(GOLDEN, "MANIFEST.json")) as f:
        return json.load(f)


def fixtures():
    # numeric arrays only: no pickled objects are loaded
    with np.load(os.path.join(GOLDEN, "rmat10_expected.npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def records():
    rec = np.fromfile(os.path.join(GOLDEN, "rmat10.bin"), dtype="<u4").reshape(-1, 2)
    return np.ascontiguousarray(rec[:, 0]), np.ascontiguousarray(rec[:, 1])


@pytest.mark.parametrize("name", ["rmat10.bin", "rmat10_expected.npz"])
def test_fixture_hash(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        assert hashlib.sha256(f.read()).hexdigest() == manifest()["sha256"][name]


def test_generator_reproduces_records():
    g = manifest()["graph"]
    src, dst = O.gen_rmat(g["scale"], g["edge_factor"] << g["scale"], seed=g["seed"])
    s, d = records()
    assert src.size == g["records"]
    assert np.array_equal(src, s) and np.array_equal(dst, d)


def test_oracle_reproduces_fixtures():
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_fixtures",
                                                  os.path.join(GOLDEN, "make_fixtures.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    now = mk.expected(*records())
    want = fixtures()
    assert sorted(now) == sorted(want)
    for k in want:
        if want[k].dtype.kind == "f":
            np.testing.assert_allclose(now[k], want[k], rtol=1e-12, atol=0, err_msg=k)
        else:
            assert np.array_equal(now[k], want[k]), k


def test_fixture_invariants():
    """What the fixtures must satisfy whatever computed them."""
    f = fixtures()
    s, d = records()
    assert np.array_equal(f["csr_ptr"], np.concatenate([[0], np.cumsum(np.bincount(s, minlength=N))]))
    assert np.array_equal(f["csc_ptr"], np.concatenate([[0], np.cumsum(np.bincount(d, minlength=N))]))
    assert int(f["u_csr_ptr"][-1]) == 2 * s.size
    assert f["grid_ptr"][-1] == s.size
    # BFS levels and WCC labels against scipy's graph routines (no oracle code)
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components, shortest_path

    a = sp.coo_matrix((np.ones(s.size), (s, d)), shape=(N, N)).tocsr()
    for root, lv in zip(f["bfs_roots"], f["bfs_levels"]):
        hops = shortest_path(a, directed=False, unweighted=True, indices=int(root))
        want = np.where(np.isfinite(hops), hops + 1, 0).astype(np.uint32)
        assert np.array_equal(lv, want), int(root)
    ncomp, comp = connected_components(a, directed=False)
    first = np.full(ncomp, N, np.int64)
    np.minimum.at(first, comp, np.arange(N))
    assert np.array_equal(f["wcc_labels"], first[comp])
    # PageRank with dangling mass dropped: the ranks sum to at most 1
    assert 0.0 < f["pagerank"].sum() <= 1.0 + 1e-12


@pytest.mark.gpu
def test_libeg_matches_fixtures():
    import everythinggraph_amd as eg

    f = fixtures()
    path = os.path.join(GOLDEN, "rmat10.bin")
    print("golden: directed layouts and PageRank", flush=True)
    with eg.Graph(N, eg.MODE_ADJ_SORT) as g:
        g.load_file(path)
        ptr, col, _ = g.csr()
        cp, row, _ = g.csc()
        assert np.array_equal(ptr, f["csr_ptr"]) and np.array_equal(col, f["csr_col"])
        assert np.array_equal(cp, f["csc_ptr"]) and np.array_equal(row, f["csc_row"])
        for binned in (0, 1):
            g.tune_set("pr_binned", binned)
            r, _, _ = g.pagerank(0.85, 20)
            err = np.abs(r.astype(np.float64) - f["pagerank"])
            assert err.sum() <= 1e-6 and err.max() <= 1e-6 * f["pagerank"].max(), binned
    print("golden: grid P=4", flush=True)
    with eg.Graph(N, eg.MODE_GRID_SORT, P=4) as g:
        g.load_file(path)
        gp, gs, gd, _ = g.grid()
        assert np.array_equal(gp, f["grid_ptr"])
        assert np.array_equal(gs, f["grid_src"]) and np.array_equal(gd, f["grid_dst"])
    print("golden: -u layout, BFS, WCC", flush=True)
    with eg.Graph(N, eg.MODE_ADJ_SORT, undirected=True) as g:
        g.load_file(path)
        ptr, col, _ = g.csr()
        assert np.array_equal(ptr, f["u_csr_ptr"]) and np.array_equal(col, f["u_csr_col"])
        for root, lv in zip(f["bfs_roots"], f["bfs_levels"]):
            dist, _, _ = g.bfs(int(root))
            assert np.array_equal(dist, lv), int(root)
        out, _ = g.bfs_batch(f["bfs_roots"])
        assert np.array_equal(out, f["bfs_levels"])
        lab, _ = g.wcc()
        assert np.array_equal(lab, f["wcc_labels"])
