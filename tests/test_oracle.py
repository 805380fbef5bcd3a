"""Pins the CPU oracle (oracle/eg_oracle.c) against independent routines and
closed forms, so that the oracle can serve as the checker of the device path.

Each oracle function is checked by something that does not share its code:
numpy stable argsort / lexsort for the layout builders, scipy.sparse products
for PageRank / SpMV / PPR, scipy.sparse.csgraph shortest paths and connected
components for BFS / WCC, brute-force transitive closure for the directed
min-label fixpoint, and closed forms (cycles, complete graphs, stars, paths).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

import eg_oracle_py as O


# ---------------------------------------------------------------------------
# small graph zoo
# ---------------------------------------------------------------------------
def rand_graph(n, m, seed, hub=None, loops=True):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, m, dtype=np.uint32)
    d = rng.integers(0, n, m, dtype=np.uint32)
    if hub is not None:
        k = m // 3
        s[:k] = hub
    if not loops:
        keep = s != d
        s, d = s[keep], d[keep]
    return s.astype(np.uint32), d.astype(np.uint32)


def zoo():
    out = []
    out.append(("empty", 5, np.zeros(0, np.uint32), np.zeros(0, np.uint32)))
    out.append(("path", 6, np.arange(5, dtype=np.uint32), np.arange(1, 6, dtype=np.uint32)))
    out.append(("self_dup", 4, np.array([0, 0, 1, 1, 3, 3], np.uint32),
                np.array([0, 1, 2, 2, 3, 0], np.uint32)))
    out.append(("isolated_tail", 50, np.array([3, 7, 3, 9], np.uint32),
                np.array([7, 3, 9, 3], np.uint32)))
    s, d = rand_graph(300, 5000, 1, hub=17)
    out.append(("hub", 300, s, d))
    s, d = O.gen_rmat(10, 16 * 1024, seed=1)
    out.append(("rmat10", 1024, s, d))
    return out


# ---------------------------------------------------------------------------
# generator (no reference counterpart): distribution and determinism
# ---------------------------------------------------------------------------
def test_rmat_quadrant_frequencies():
    scale, m = 12, 200000
    s, d = O.gen_rmat(scale, m, seed=1)
    assert s.max() < (1 << scale) and d.max() < (1 << scale)
    # every level is an independent quadrant draw: top bits follow (a,b,c,d)
    top_s = (s >> (scale - 1)) & 1
    top_d = (d >> (scale - 1)) & 1
    freq = np.bincount(top_s * 2 + top_d, minlength=4) / m
    np.testing.assert_allclose(freq, [0.57, 0.19, 0.19, 0.05], atol=0.005)
    low_s, low_d = s & 1, d & 1
    freq = np.bincount(low_s * 2 + low_d, minlength=4) / m
    np.testing.assert_allclose(freq, [0.57, 0.19, 0.19, 0.05], atol=0.005)


def test_generator_determinism_and_seeds():
    a = O.gen_rmat(10, 5000, seed=7)
    b = O.gen_rmat(10, 5000, seed=7)
    c = O.gen_rmat(10, 5000, seed=8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])
    # a prefix of a longer stream is the shorter stream (counter-based)
    p = O.gen_rmat(10, 1000, seed=7)
    assert np.array_equal(p[0], a[0][:1000]) and np.array_equal(p[1], a[1][:1000])


def test_er_uniform_and_weights():
    n, m = 1000, 400000
    s, d = O.gen_er(n, m, seed=3)
    assert s.max() < n and d.max() < n
    hist = np.bincount(s, minlength=n)
    assert abs(hist.mean() - m / n) < 1e-9 and hist.std() < 3 * np.sqrt(m / n)
    w = O.gen_weights(m, seed=3)
    assert w.min() > 0 and w.max() <= 1.0
    k = w.astype(np.float64) * 2.0**24
    assert np.array_equal(k, np.round(k))  # exact multiples of 2^-24
    assert abs(w.mean() - 0.5) < 0.01


def test_symmetrize_order():
    s = np.array([0, 1, 2], np.uint32)
    d = np.array([1, 2, 0], np.uint32)
    w = np.array([0.5, 0.25, 1.0], np.float32)
    os_, od, ow = O.symmetrize(s, d, w)
    assert list(os_) == [0, 1, 2, 1, 2, 0]
    assert list(od) == [1, 2, 0, 0, 1, 2]
    assert list(ow) == [0.5, 0.25, 1.0, 0.5, 0.25, 1.0]


# ---------------------------------------------------------------------------
# partition geometry (init_all.c:29-60)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,Pp", [(10, 3), (16, 4), (7, 7), (5, 8), (1000, 16), (65536, 25),
                                  (1 << 22, 1638)])
def test_partition_geometry(n, Pp):
    vs = np.unique(np.linspace(0, n - 1, min(n, 20000)).astype(np.int64))
    ids = np.array([O.partition_id(n, Pp, int(v)) for v in vs])
    assert np.all(np.diff(ids) >= 0) and ids[0] == 0 and ids[-1] <= Pp - 1
    if n <= 20000:
        sizes = np.bincount(ids, minlength=Pp)
        if n >= Pp:
            assert sizes.max() - sizes.min() <= 1 and sizes.sum() == n
            # the first n % P partitions hold the larger size
            big = n % Pp
            assert np.all(sizes[:big] == n // Pp + 1) and np.all(sizes[big:] == n // Pp)


def test_default_P():
    # random.c:69-71
    assert O.default_P(1 << 16) == 25
    assert O.default_P(1 << 22) == 1638
    assert O.default_P(1 << 24) == 256
    assert O.default_P(100) == 4
    assert O.default_P(1 << 30) == 256  # widened arithmetic (R13)


# ---------------------------------------------------------------------------
# builders vs numpy stable sorts
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name,n,s,d", zoo(), ids=[z[0] for z in zoo()])
def test_csr_csc_vs_numpy(name, n, s, d):
    w = (np.arange(s.size, dtype=np.float32) + 1) / 4096
    ptr, col, ow = O.build_csr(n, s, d, w)
    perm = np.argsort(s, kind="stable")
    assert np.array_equal(col, d[perm]) and np.array_equal(ow, w[perm])
    ref_ptr = np.concatenate([[0], np.cumsum(np.bincount(s, minlength=n))])
    assert np.array_equal(ptr, ref_ptr)
    cp, row, iw = O.build_csc(n, ptr, col, ow)
    rowof = np.repeat(np.arange(n, dtype=np.uint32), np.diff(ptr).astype(np.int64))
    perm2 = np.argsort(col, kind="stable")
    assert np.array_equal(row, rowof[perm2]) and np.array_equal(iw, ow[perm2])
    assert np.array_equal(cp, np.concatenate([[0], np.cumsum(np.bincount(d, minlength=n))]))


@pytest.mark.parametrize("n", [1 << 11, 1 << 20, (1 << 24) + 3])
def test_parallel_sort_path_vs_numpy(n):
    """From 2^22 records the oracle sorts with a multi-threaded LSD radix sort
    (11-bit digits; 1, 2 and 3 passes here): same CSR / CSC as numpy's stable
    argsort, including skewed keys and keys past 2^22."""
    m = (1 << 22) + 17
    s, d = O.gen_rmat(11, m, seed=9)
    rng = np.random.default_rng(n)
    s = (s.astype(np.uint64) * (n // 2048) + rng.integers(0, n // 2048 or 1, m)).astype(np.uint32)
    s = np.minimum(s, n - 1).astype(np.uint32)
    d = rng.integers(0, n, m).astype(np.uint32)
    ptr, col, _ = O.build_csr(n, s, d)
    perm = np.argsort(s, kind="stable")
    assert np.array_equal(col, d[perm])
    assert np.array_equal(ptr, np.concatenate([[0], np.cumsum(np.bincount(s, minlength=n))]))
    cp, row, _ = O.build_csc(n, ptr, col)
    rowof = np.repeat(np.arange(n, dtype=np.uint32), np.diff(ptr).astype(np.int64))
    assert np.array_equal(row, rowof[np.argsort(col, kind="stable")])
    assert np.array_equal(cp, np.concatenate([[0], np.cumsum(np.bincount(d, minlength=n))]))


@pytest.mark.parametrize("name,n,s,d", zoo(), ids=[z[0] for z in zoo()])
@pytest.mark.parametrize("Pp", [1, 3, 16])
def test_grid_vs_numpy(name, n, s, d, Pp):
    w = (np.arange(s.size, dtype=np.float32) + 1) / 4096
    part = np.array([O.partition_id(n, Pp, v) for v in range(n)], np.int64)
    for presort in (True, False):
        cp, gs, gd, gw = O.build_grid(n, Pp, s, d, w, presort=presort)
        order = np.argsort(s, kind="stable") if presort else np.arange(s.size)
        cell = part[s[order]] * Pp + part[d[order]]
        perm = order[np.argsort(cell, kind="stable")]
        assert np.array_equal(gs, s[perm]) and np.array_equal(gd, d[perm])
        assert np.array_equal(gw, w[perm])
        cnt = np.bincount(part[s] * Pp + part[d], minlength=Pp * Pp)
        assert np.array_equal(cp, np.concatenate([[0], np.cumsum(cnt)]))


@pytest.mark.parametrize("name,n,s,d", zoo(), ids=[z[0] for z in zoo()])
def test_edgelist_vs_numpy(name, n, s, d):
    w = (np.arange(s.size, dtype=np.float32) + 1) / 4096
    for mode in (7, 8, 9):
        os_, od, ow = O.build_edgelist(n, s, d, w, mode)
        if mode == 9:
            perm = np.arange(s.size)
        elif mode == 7:
            perm = np.argsort(s, kind="stable")
        else:
            perm = np.lexsort((np.arange(s.size), d, s))
        assert np.array_equal(os_, s[perm]) and np.array_equal(od, d[perm])
        assert np.array_equal(ow, w[perm])


# ---------------------------------------------------------------------------
# PageRank / SpMV / PPR vs scipy
# ---------------------------------------------------------------------------
def scipy_pr(n, s, d, damping, iters, init, uniform=False):
    A = sp.csr_matrix((np.ones(s.size), (s.astype(np.int64), d.astype(np.int64))), shape=(n, n))
    deg = np.asarray(A.sum(axis=1)).ravel()
    r = np.full(n, init)
    for _ in range(iters):
        contrib = np.where(deg > 0, r / np.maximum(deg, 1), 0.0)
        dang = r[deg == 0].sum()
        r = (1 - damping) / n + damping * (A.T @ contrib) + (damping * dang / n if uniform else 0)
    return r


@pytest.mark.parametrize("name,n,s,d", zoo()[1:], ids=[z[0] for z in zoo()[1:]])
@pytest.mark.parametrize("uniform", [False, True])
def test_pagerank_vs_scipy(name, n, s, d, uniform):
    ptr, col, _ = O.build_csr(n, s, d)
    cp, row, _ = O.build_csc(n, ptr, col)
    deg = O.out_degree(n, s)
    r, dh, lh = O.pagerank_pull(n, cp, row, deg, 0.85, 20, dangling_uniform=uniform)
    ref = scipy_pr(n, s, d, 0.85, 20, 1.0 / n, uniform)
    np.testing.assert_allclose(r, ref, rtol=1e-12, atol=1e-15)
    assert lh[-1] < lh[0] or lh[0] == 0
    # dangling mass of the input iterate of iteration 0 is init * #sinks
    assert dh[0] == pytest.approx((deg == 0).sum() / n, rel=1e-12)
    if uniform and name != "empty":
        assert r.sum() == pytest.approx(1.0, rel=1e-12)  # mass is conserved
    Pp = 3
    gcp, gs, gd, _ = O.build_grid(n, Pp, s, d)
    rg = O.pagerank_grid(n, Pp, gcp, gs, gd, deg, 0.85, 20)
    if not uniform:
        np.testing.assert_allclose(rg, r, rtol=1e-12, atol=1e-15)


def test_pagerank_closed_forms():
    # directed cycle: every vertex keeps 1/n
    n = 7
    s = np.arange(n, dtype=np.uint32)
    d = ((s + 1) % n).astype(np.uint32)
    ptr, col, _ = O.build_csr(n, s, d)
    cp, row, _ = O.build_csc(n, ptr, col)
    r, _, lh = O.pagerank_pull(n, cp, row, O.out_degree(n, s), 0.85, 10)
    np.testing.assert_allclose(r, 1.0 / n, rtol=1e-14)
    assert np.all(lh < 1e-15)
    # complete digraph without loops: stays uniform
    n = 5
    s, d = np.meshgrid(np.arange(n), np.arange(n))
    keep = s != d
    s, d = s[keep].astype(np.uint32), d[keep].astype(np.uint32)
    ptr, col, _ = O.build_csr(n, s, d)
    cp, row, _ = O.build_csc(n, ptr, col)
    r, _, _ = O.pagerank_pull(n, cp, row, O.out_degree(n, s), 0.85, 10)
    np.testing.assert_allclose(r, 1.0 / n, rtol=1e-14)
    # star into a sink (leaves -> 0): one iteration from r0 = 1/n
    n = 6
    s = np.arange(1, n, dtype=np.uint32)
    d = np.zeros(n - 1, np.uint32)
    ptr, col, _ = O.build_csr(n, s, d)
    cp, row, _ = O.build_csc(n, ptr, col)
    r, dh, _ = O.pagerank_pull(n, cp, row, O.out_degree(n, s), 0.85, 1)
    assert r[0] == pytest.approx(0.15 / n + 0.85 * (n - 1) / n, rel=1e-14)
    np.testing.assert_allclose(r[1:], 0.15 / n, rtol=1e-14)
    assert dh[0] == pytest.approx(1.0 / n, rel=1e-14)


@pytest.mark.parametrize("name,n,s,d", zoo()[1:], ids=[z[0] for z in zoo()[1:]])
def test_spmv_vs_scipy(name, n, s, d):
    w = O.gen_weights(s.size, seed=5)
    ptr, col, ow = O.build_csr(n, s, d, w)
    cp, row, iw = O.build_csc(n, ptr, col, ow)
    x = np.linspace(-1, 1, n)
    y = O.spmv_pull(n, cp, row, iw, x)
    A = sp.csr_matrix((w.astype(np.float64), (s.astype(np.int64), d.astype(np.int64))), shape=(n, n))
    np.testing.assert_allclose(y, A.T @ x, rtol=1e-12, atol=1e-12)
    gcp, gs, gd, gw = O.build_grid(n, 4, s, d, w)
    np.testing.assert_allclose(O.spmv_grid(n, 4, gcp, gs, gd, gw, x), A.T @ x, rtol=1e-12,
                               atol=1e-12)
    # identity matrix: y = x
    I = np.arange(n, dtype=np.uint32)
    ip, ic, iw2 = O.build_csr(n, I, I, np.ones(n, np.float32))
    np.testing.assert_array_equal(O.spmv_pull(n, ip, ic, iw2, x), x)


@pytest.mark.parametrize("name,n,s,d", zoo()[3:], ids=[z[0] for z in zoo()[3:]])
def test_ppr_vs_scipy(name, n, s, d):
    ptr, col, _ = O.build_csr(n, s, d)
    cp, row, _ = O.build_csc(n, ptr, col)
    deg = O.out_degree(n, s)
    seeds = np.unique(s)[:3]
    out = O.ppr(n, cp, row, deg, seeds, 0.15, 30)
    A = sp.csr_matrix((np.ones(s.size), (s.astype(np.int64), d.astype(np.int64))), shape=(n, n))
    p = np.zeros(n)
    p[seeds] = 1.0 / seeds.size
    x = p.copy()
    dg = np.asarray(A.sum(axis=1)).ravel()
    for _ in range(30):
        x = 0.15 * p + 0.85 * (A.T @ np.where(dg > 0, x / np.maximum(dg, 1), 0.0))
    np.testing.assert_allclose(out, x, rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------------------
# BFS vs scipy shortest paths
# ---------------------------------------------------------------------------
def scipy_levels(n, s, d, root):
    A = sp.csr_matrix((np.ones(s.size), (s.astype(np.int64), d.astype(np.int64))), shape=(n, n))
    dist = csg.shortest_path(A, unweighted=True, indices=[root], directed=True)[0]
    out = np.where(np.isinf(dist), 0, dist + 1)
    return out.astype(np.uint32)


@pytest.mark.parametrize("name,n,s,d", zoo()[1:], ids=[z[0] for z in zoo()[1:]])
@pytest.mark.parametrize("undirected", [False, True])
def test_bfs_vs_scipy(name, n, s, d, undirected):
    if undirected:
        s, d, _ = O.symmetrize(s, d)
    ptr, col, _ = O.build_csr(n, s, d)
    cp, row, _ = O.build_csc(n, ptr, col)
    gcp, gs, gd, _ = O.build_grid(n, 4, s, d)
    roots = sorted(set([0, int(s[0]) if s.size else 0, n - 1]))
    for root in roots:
        ref = scipy_levels(n, s, d, root)
        for start_pull in (False, True):
            for switch in (False, True):
                dist, lv, ls = O.bfs(n, s.size, ptr, col, cp, row, root, start_pull, switch)
                assert np.array_equal(dist, ref)
                assert lv == ref.max()
                assert np.array_equal(ls, np.bincount(ref, minlength=lv + 1)[1:lv + 1])
        assert np.array_equal(O.bfs_grid(n, 4, gcp, gs, gd, root)[0], ref)
        assert np.array_equal(O.bfs_edgelist(n, s, d, root)[0], ref)


def test_bfs_closed_forms():
    n = 9
    s = np.arange(n - 1, dtype=np.uint32)
    d = s + 1
    ptr, col, _ = O.build_csr(n, s, d)
    cp, row, _ = O.build_csc(n, ptr, col)
    dist, lv, _ = O.bfs(n, s.size, ptr, col, cp, row, 0)
    assert list(dist) == list(range(1, n + 1)) and lv == n
    dist, _, _ = O.bfs(n, s.size, ptr, col, cp, row, 4)
    assert list(dist) == [0] * 4 + [1, 2, 3, 4, 5]


def test_parent_checker_catches_planted_errors():
    n, (s, d) = 1024, O.gen_rmat(10, 8192, seed=4)
    s, d, _ = O.symmetrize(s, d)
    ptr, col, _ = O.build_csr(n, s, d)
    root = int(s[0])
    dist = scipy_levels(n, s, d, root)
    # a legal parent: the first out-neighbour (symmetric) one level up
    parent = np.full(n, n, np.uint32)
    parent[root] = root
    for v in range(n):
        if dist[v] > 1:
            nb = col[ptr[v]:ptr[v + 1]]
            parent[v] = nb[dist[nb] == dist[v] - 1][0]
    assert O.check_bfs_parents(n, ptr, col, root, dist, parent) == 0
    reached = np.nonzero(dist > 1)[0]
    bad = parent.copy()
    bad[reached[0]] = reached[0]  # self parent
    assert O.check_bfs_parents(n, ptr, col, root, dist, bad) > 0
    bad = parent.copy()
    bad[root] = n  # root not its own parent
    assert O.check_bfs_parents(n, ptr, col, root, dist, bad) > 0
    bad_dist = dist.copy()
    bad_dist[reached[0]] += 1  # level off by one
    assert O.check_bfs_parents(n, ptr, col, root, bad_dist, parent) > 0
    # a parent one level up that is not an in-neighbour
    for v in reached:
        nb = set(col[ptr[v]:ptr[v + 1]].tolist())
        up = [u for u in np.nonzero(dist == dist[v] - 1)[0] if u not in nb]
        if up:
            bad = parent.copy()
            bad[v] = up[0]
            assert O.check_bfs_parents(n, ptr, col, root, dist, bad) == 1
            break
    else:
        raise AssertionError("no vertex with a non-adjacent vertex one level up")


# ---------------------------------------------------------------------------
# full-size checkers (row content from the records, rows of selected
# vertices, BFS level invariants) vs numpy / scipy
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name,n,s,d", zoo(), ids=[z[0] for z in zoo()])
@pytest.mark.parametrize("both", [0, 1])
def test_row_sums_and_rows_of_vs_numpy(name, n, s, d, both):
    ss, dd = (np.concatenate([s, d]), np.concatenate([d, s])) if both else (s, d)
    cnt, sm, xr = O.record_row_sums(n, s, d, both)
    assert np.array_equal(cnt, np.bincount(ss, minlength=n))
    ref_sum = np.zeros(n, np.uint64)
    np.add.at(ref_sum, ss.astype(np.int64), dd.astype(np.uint64))
    ref_xr = np.zeros(n, np.uint32)
    np.bitwise_xor.at(ref_xr, ss.astype(np.int64), dd)
    assert np.array_equal(sm, ref_sum) and np.array_equal(xr, ref_xr)
    # the same content read from the rows of a CSR (numpy's stable sort)
    perm = np.argsort(ss, kind="stable")
    ptr = np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint64)
    col = dd[perm]
    csum, cxr = O.csr_row_sums(n, ptr, col)
    assert np.array_equal(csum, ref_sum) and np.array_equal(cxr, ref_xr)
    # one entry moved to the next row changes both checks
    if col.size > 1 and cnt[ss[perm[0]]] > 0:
        bad = ptr.copy()
        r = int(ss[perm[0]])
        bad[r + 1] -= 1
        bsum, bxr = O.csr_row_sums(n, bad, col)
        assert not (np.array_equal(bsum, ref_sum) and np.array_equal(bxr, ref_xr))
    # rows of selected vertices in stored order, hubs included
    rng = np.random.default_rng(3)
    sel = np.unique(np.concatenate([rng.integers(0, n, 7), [int(np.argmax(cnt))]])).astype(np.uint32)
    rp, rc = O.rows_of(n, s, d, both, sel)
    for j, v in enumerate(sel):
        assert np.array_equal(rc[rp[j]:rp[j + 1]], col[ptr[v]:ptr[v + 1]]), (name, v)


@pytest.mark.parametrize("undirected", [False, True])
def test_bfs_level_checker_vs_scipy_and_planted_errors(undirected):
    n, (s, d) = 1024, O.gen_rmat(10, 8192, seed=6)
    ss, dd = (O.symmetrize(s, d)[:2] if undirected else (s, d))
    root = int(s[0])
    dist = scipy_levels(n, ss, dd, root)
    assert O.check_bfs_levels(n, s, d, int(undirected), root, dist) == 0
    reached = np.nonzero(dist > 1)[0]
    deep = reached[np.argmax(dist[reached])]
    for plant in range(5):
        bad = dist.copy()
        if plant == 0:
            bad[reached[0]] += 1      # a level one too deep
        elif plant == 1:
            bad[deep] -= 1            # one too shallow: no neighbour a level up
        elif plant == 2:
            bad[reached[0]] = 0       # a reachable vertex left unreached
        elif plant == 3:
            bad[dist > 0] += 1        # every level shifted: root is not 1
        else:
            bad = scipy_levels(n, ss, dd, int(reached[0]))  # another root's levels
        assert O.check_bfs_levels(n, s, d, int(undirected), root, bad) > 0, plant


# ---------------------------------------------------------------------------
# WCC vs scipy connected components / brute force
# ---------------------------------------------------------------------------
def min_label(n, s, d):
    A = sp.csr_matrix((np.ones(s.size), (s.astype(np.int64), d.astype(np.int64))), shape=(n, n))
    _, comp = csg.connected_components(A, directed=True, connection="weak")
    mins = np.full(comp.max() + 1, n)
    np.minimum.at(mins, comp, np.arange(n))
    return mins[comp].astype(np.int64)


@pytest.mark.parametrize("name,n,s,d", zoo(), ids=[z[0] for z in zoo()])
def test_wcc_vs_scipy(name, n, s, d):
    ref = min_label(n, s, d)
    lab, _ = O.wcc_edges(n, s, d)
    assert np.array_equal(lab, ref)
    us, ud, _ = O.symmetrize(s, d)
    ptr, col, _ = O.build_csr(n, us, ud)
    lab2, _ = O.wcc_push(n, ptr, col)
    assert np.array_equal(lab2, ref)


def test_wcc_push_directed_is_min_ancestor():
    n, m = 60, 90
    s, d = rand_graph(n, m, 11, loops=False)
    ptr, col, _ = O.build_csr(n, s, d)
    lab, _ = O.wcc_push(n, ptr, col)
    # brute force: reach[u][v] by repeated squaring of the adjacency
    R = np.eye(n, dtype=bool)
    R[s, d] = True
    for _ in range(7):
        R = R | ((R.astype(np.int32) @ R.astype(np.int32)) > 0)
    ref = np.array([np.nonzero(R[:, v])[0].min() for v in range(n)], np.int64)
    assert np.array_equal(lab, ref)


# ---------------------------------------------------------------------------
# SSSP (sssp_pushpull.c)
# ---------------------------------------------------------------------------
def bellman_ford_np(n, s, d, w, root):
    """Brute-force Bellman-Ford over the edge list, vectorised, no shared code."""
    INF = np.uint64(O.SSSP_UNREACHED)
    dist = np.full(n, INF, np.uint64)
    dist[root] = 0
    keep = s != d
    s, d, w = s[keep].astype(np.int64), d[keep].astype(np.int64), w[keep]
    for _ in range(n + 1):
        cand = np.where(dist[s] < INF, dist[s] + w, INF)
        new = dist.copy()
        np.minimum.at(new, d, cand)
        if np.array_equal(new, dist):
            break
        dist = new
    return dist.astype(np.uint32)


@pytest.mark.parametrize("name,n,s,d", zoo()[1:], ids=[z[0] for z in zoo()[1:]])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_sssp_vs_independent(name, n, s, d, kind):
    ptr, col, _ = O.build_csr(n, s, d)
    roots = sorted({0, int(s[0]) if s.size else 0, n - 1})
    for r in roots:
        dist, _ = O.sssp(n, ptr, col, r, kind)
        w = O.sssp_weight(s, d, kind)
        assert np.array_equal(dist, bellman_ford_np(n, s, d, w, r)), (name, kind, r)
        if kind == 0:
            continue
        # library shortest paths (weights must be > 0 there: zero means no edge)
        keep = s != d
        A = sp.coo_matrix((O.sssp_weight(s[keep], d[keep], kind).astype(np.float64),
                           (s[keep], d[keep])), shape=(n, n)).tocsr()
        A.sum_duplicates()
        A.data = np.minimum.reduceat(
            O.sssp_weight(*_coo_sorted(s[keep], d[keep]), kind).astype(np.float64),
            _run_starts(*_coo_sorted(s[keep], d[keep]))) if keep.any() else A.data
        ref = csg.dijkstra(A, directed=True, indices=r)
        ref = np.where(np.isinf(ref), O.SSSP_UNREACHED, ref).astype(np.uint64)
        assert np.array_equal(dist.astype(np.uint64), ref), (name, kind, r)


def _coo_sorted(s, d):
    o = np.lexsort((d, s))
    return s[o], d[o]


def _run_starts(s, d):
    key = s.astype(np.uint64) << np.uint64(32) | d.astype(np.uint64)
    return np.flatnonzero(np.r_[True, key[1:] != key[:-1]])


def test_sssp_closed_forms():
    # path 0 -> 1 -> ... -> 9: dist[k] = sum of (i%10 + (i+1)%10) for i < k
    n = 10
    s = np.arange(9, dtype=np.uint32)
    d = s + 1
    ptr, col, _ = O.build_csr(n, s, d)
    dist, _ = O.sssp(n, ptr, col, 0, 0)
    expect = np.concatenate([[0], np.cumsum([i % 10 + (i + 1) % 10 for i in range(9)])])
    assert np.array_equal(dist, expect.astype(np.uint32))
    # unreachable vertices and an out-of-graph root
    dist, _ = O.sssp(n, ptr, col, 5, 1)
    assert np.all(dist[:5] == O.SSSP_UNREACHED) and np.array_equal(dist[5:], np.arange(5))
    # zero-weight edges (10 -> 20 has weight 0) still give shortest paths
    s = np.array([10, 20, 10], np.uint32)
    d = np.array([20, 30, 30], np.uint32)
    ptr, col, _ = O.build_csr(40, s, d)
    dist, _ = O.sssp(40, ptr, col, 10, 0)
    assert dist[20] == 0 and dist[30] == 0


def test_sssp_parent_checker_catches_planted_errors():
    n = 1 << 9
    s, d = O.gen_rmat(9, 8 << 9, seed=17)
    ptr, col, _ = O.build_csr(n, s, d)
    root = int(s[0])
    dist, _ = O.sssp(n, ptr, col, root, 0)
    # a legal tree: for every reached v, the first in-neighbour on a shortest path
    parent = np.full(n, O.SSSP_UNREACHED, np.uint32)
    parent[root] = root
    for u in range(n):
        for k in range(ptr[u], ptr[u + 1]):
            v = col[k]
            if v != root and v != u and parent[v] == O.SSSP_UNREACHED and dist[u] != O.SSSP_UNREACHED \
                    and int(dist[u]) + int(O.sssp_weight(u, v, 0)) == int(dist[v]):
                parent[v] = u
    assert O.check_sssp_parents(n, ptr, col, root, 0, dist, parent) == 0
    reached = np.flatnonzero((dist != O.SSSP_UNREACHED) & (np.arange(n) != root))
    bad = parent.copy()
    bad[reached[0]] = root if parent[reached[0]] != root else reached[1]
    bad2 = parent.copy()
    bad2[root] = reached[0]
    dist3 = dist.copy()
    dist3[reached[2]] += 1
    errs = [O.check_sssp_parents(n, ptr, col, root, 0, dist, bad),
            O.check_sssp_parents(n, ptr, col, root, 0, dist, bad2),
            O.check_sssp_parents(n, ptr, col, root, 0, dist3, parent)]
    assert all(e > 0 for e in errs), errs
