"""Parity at the benchmark's own size: R-MAT 24, edge factor 16, -u -s
(n = 16,777,216, 536,870,912 stored entries), generated and built on the
device as bench.py does, checked against the oracle and against properties
that do not depend on size.

* Layout: the degrees and the per-row sums of neighbour ids equal those of
  the oracle's generator output (a checksum of every row), row offsets are a
  prefix sum ending at m.
* PageRank (20 iterations, the binned kernels that the bench times, and the
  pull kernel): L1 error <= 1e-6 and max error <= 1e-6 of the largest rank
  against the FP64 oracle.
* SpMV with signed x: |y - y64| <= 1e-5 (|A| |x|) elementwise.
* BFS: levels bit-exact against the oracle, parents a legal BFS tree, the
  Graph500 edge count; multi-source BFS over 256 roots (the batched
  configuration) equal to single-source, four roots equal to the oracle.
* WCC labels bit-exact; PPR, 64 queries batched as one SpMM (the batched
  configuration) and single-vector, L1 <= 2e-6 against each other and, for
  four queries, <= 1e-6 against the FP64 oracle.
"""
import numpy as np
import pytest

import eg_oracle_py as O
import everythinggraph_amd as eg

pytestmark = pytest.mark.gpu

SCALE, EF = 24, 16
N = 1 << SCALE


def say(msg):
    print(f"[full size R-MAT {SCALE}] {msg}", flush=True)


@pytest.fixture(scope="module")
def big():
    say("generating and building on the device")
    g = eg.Graph(N, eg.MODE_ADJ_SORT, undirected=True, symmetric=True)
    g.generate(eg.rmat_params(SCALE, EF, seed=1))
    ptr, col, _ = g.csr()
    deg = np.diff(ptr).astype(np.uint32)
    yield g, ptr, col, deg
    g.close()


def test_full_size_layout(big):
    g, ptr, col, deg = big
    m = g.info.m
    assert m == 2 * (EF << SCALE) and ptr[0] == 0 and int(ptr[-1]) == m
    assert np.all(ptr[1:] >= ptr[:-1]) and int(col.max()) < N
    say("oracle generator")
    s, d = O.gen_rmat(SCALE, EF << SCALE, seed=1)
    ref_deg = np.bincount(s, minlength=N) + np.bincount(d, minlength=N)
    assert np.array_equal(deg, ref_deg)
    ref_sum = (np.bincount(s, weights=d.astype(np.float64), minlength=N)
               + np.bincount(d, weights=s.astype(np.float64), minlength=N))
    del s, d
    cs = np.zeros(m + 1, np.uint64)
    np.cumsum(col, dtype=np.uint64, out=cs[1:])
    row_sum = cs[ptr[1:].astype(np.int64)] - cs[ptr[:-1].astype(np.int64)]
    del cs
    assert np.array_equal(row_sum, ref_sum.astype(np.uint64))


@pytest.mark.parametrize("binned", [1, 0])
def test_full_size_pagerank(big, binned):
    """The binned kernels (what the bench times) and the pull kernel."""
    g, ptr, col, deg = big
    say(f"pagerank, 20 iterations, pr_binned={binned}")
    g.tune_set("pr_binned", binned)
    r, _, _ = g.pagerank(0.85, 20)
    g.tune_set("pr_binned", -1)
    ref, _, _ = O.pagerank_pull(N, ptr, col, deg, 0.85, 20)
    err = np.abs(r.astype(np.float64) - ref)
    assert err.sum() <= 1e-6
    assert err.max() <= 1e-6 * ref.max()


def test_full_size_spmv(big):
    g, ptr, col, deg = big
    say("spmv")
    x = np.random.default_rng(5).uniform(-1, 1, N).astype(np.float32)
    y = g.spmv(x)
    ref = O.spmv_pull(N, ptr, col, None, x.astype(np.float64))
    absref = O.spmv_pull(N, ptr, col, None, np.abs(x).astype(np.float64))
    assert np.all(np.abs(y.astype(np.float64) - ref) <= 1e-5 * absref + 1e-30)


def test_full_size_bfs(big):
    g, ptr, col, deg = big
    for root in eg.sample_roots(deg, 2, seed=2):
        say(f"bfs root {int(root)}")
        ref, lv, _ = O.bfs(N, int(ptr[-1]), ptr, col, ptr, col, int(root))
        dist, par, st = g.bfs(int(root))
        assert np.array_equal(dist, ref)
        assert st.levels == lv
        assert O.check_bfs_parents(N, ptr, col, int(root), dist, par) == 0
        reached = ref != 0
        assert st.reached == reached.sum()
        assert st.edges_traversed == int(deg[reached].sum()) // 2


@pytest.mark.timeout(600)
def test_full_size_multi_source_bfs(big):
    """The batched configuration: 256 BFS sources in one multi-source call;
    every lane equals the single-source traversal from its root, and four
    lanes equal the oracle."""
    g, ptr, col, deg = big
    roots = eg.sample_roots(deg, 256, seed=3)
    say("multi-source bfs, 256 roots")
    out, _ = g.bfs_batch(roots)
    for i, root in enumerate(roots):
        dist, _, _ = g.bfs(int(root), parents=False)
        assert np.array_equal(out[i], dist), i
    for i in (0, 85, 170, 255):
        ref, _, _ = O.bfs(N, int(ptr[-1]), ptr, col, ptr, col, int(roots[i]))
        assert np.array_equal(out[i], ref), i
    del out


def test_full_size_wcc(big):
    g, ptr, col, deg = big
    say("wcc")
    lab, _ = g.wcc()
    ref, _ = O.wcc_push(N, ptr, col)
    assert np.array_equal(lab, ref)


@pytest.mark.timeout(600)
def test_full_size_ppr(big):
    """The batched configuration: 64 personalized PageRank vectors (1-8 seeds,
    alpha 0.15, 30 iterations) as one SpMM; four against the FP64 oracle
    within 1e-6 L1, and all 64 against the single-vector path within 2e-6."""
    g, ptr, col, deg = big
    rng = np.random.default_rng(3)
    seeds = [eg.sample_roots(deg, int(rng.integers(1, 9)), seed=1000 + q) for q in range(64)]
    say("ppr, 64 queries batched, 30 iterations")
    out, _ = g.ppr_batch(seeds, 0.15, 30)
    for q, sd in enumerate(seeds):
        one, it, _ = g.ppr(sd, 0.15, tol=0.0, max_iter=30)
        assert it == 30
        # each is within 1e-6 of the oracle, so within 2e-6 of each other
        assert np.abs(one.astype(np.float64) - out[q]).sum() <= 2e-6, q
        if q % 21 == 0:
            say(f"ppr query {q} against the oracle")
            ref = O.ppr(N, ptr, col, deg, sd, 0.15, 30)
            assert np.abs(out[q].astype(np.float64) - ref).sum() <= 1e-6, q
            assert np.abs(one.astype(np.float64) - ref).sum() <= 1e-6, q
