"""Adjacency layouts with 64-bit row offsets (2^32 stored entries or more).

From 2^32 stored entries on, the adjacency modes build the CSR / CSC from
stably sorted parts of at most 2^31 records, placed by 64-bit row offsets;
BFS runs on them.  Small graphs are built the same way with forced small
parts ("wide_part") and must equal the oracle's CSR / CSC bit for bit, their
BFS levels and level counts the oracle's queue BFS.  The full-size case,
R-MAT 27 with -u -s (2^31 records, 2^32 stored entries), is checked against
its records: the degree, column sum and column xor of every row, the exact
rows of sampled vertices (hubs included, whose rows span both parts), and
the single-source invariants of BFS levels and parents.
"""
import time

import numpy as np
import pytest

import eg_oracle_py as O
import everythinggraph_amd as eg

pytestmark = pytest.mark.gpu


def stored(s0, d0, w0, undirected):
    if not undirected:
        return s0, d0, w0
    return O.symmetrize(s0, d0, w0)


def check_against_oracle(g, n, s0, d0, w0, undirected, symmetric, roots):
    s, d, w = stored(s0, d0, w0, undirected)
    info = g.info
    assert info.m == s.size and info.has_csr == 1 and info.has_csc == int(not symmetric)
    ptr, col, ow = g.csr()
    rp, rc, rw = O.build_csr(n, s, d, w)
    assert np.array_equal(ptr, rp) and np.array_equal(col, rc)
    if w is not None:
        assert np.array_equal(ow, rw)
    out_deg, in_deg = g.degrees()
    assert np.array_equal(out_deg, np.diff(rp))
    if symmetric:
        cp, row = rp, rc
        assert np.array_equal(in_deg, out_deg)
    else:
        cp, row, cw = g.csc()
        ep, er, ew = O.build_csc(n, rp, rc, rw)
        assert np.array_equal(cp, ep) and np.array_equal(row, er)
        if w is not None:
            assert np.array_equal(cw, ew)
        assert np.array_equal(in_deg, np.diff(ep))
        cp, row = ep, er
    for r in roots:
        for start_pull, switch in ((False, True), (True, True), (False, False)):
            dist, par, st = g.bfs(int(r), start_pull=start_pull, switch=switch)
            ref, lv, _ = O.bfs(n, s.size, rp, rc, cp, row, int(r), start_pull, switch)
            assert np.array_equal(dist, ref), (r, start_pull, switch)
            assert st.levels == lv and st.reached == int((ref > 0).sum())
            assert O.check_bfs_parents(n, rp, rc, int(r), dist, par) == 0
            if not switch and not start_pull:
                assert st.pull_levels == 0


@pytest.mark.parametrize("part", [10, 13])
@pytest.mark.parametrize("undirected,symmetric", [(False, False), (True, False), (True, True)])
@pytest.mark.parametrize("weighted", [False, True])
def test_forced_parts_equal_oracle(part, undirected, symmetric, weighted):
    n = 1 << 11
    s0, d0 = O.gen_rmat(11, 16 << 11, seed=31 + part)
    w0 = O.gen_weights(s0.size, seed=5) if weighted else None
    with eg.Graph(n, eg.MODE_ADJ_SORT, undirected=undirected, symmetric=symmetric,
                  weighted=weighted) as g:
        g.tune_set("wide_part", part)
        print(f"wide build part=2^{part} u={undirected} s={symmetric} w={weighted}", flush=True)
        g.load_edges(s0, d0, w0)
        roots = eg.sample_roots(g.degrees()[0], 3, seed=part)
        check_against_oracle(g, n, s0, d0, w0, undirected, symmetric, roots)
        # the 32-bit-offset kernels refuse the layout
        with pytest.raises(eg.EgError) as e:
            g.pagerank(iterations=2)
        assert e.value.code == eg.EG_ENOTSUP


def test_count_mode_and_edge_cases():
    """Mode 4 builds the same layout; a hub row spread over many parts; an
    empty graph; records in descending order."""
    n = 3000
    rng = np.random.default_rng(4)
    s0 = np.concatenate([np.zeros(5000, np.uint32), rng.integers(0, n, 4000).astype(np.uint32)])
    d0 = rng.integers(0, n, s0.size).astype(np.uint32)
    for mode, (a, b) in ((eg.MODE_ADJ_COUNT, (s0, d0)),
                         (eg.MODE_ADJ_SORT, (np.sort(s0)[::-1].copy(), d0))):
        with eg.Graph(n, mode) as g:
            g.tune_set("wide_part", 10)  # row 0 spans at least five parts
            print(f"wide edge case mode={mode}", flush=True)
            g.load_edges(a, b)
            check_against_oracle(g, n, a, b, None, False, False, [0, int(b[7])])
    with eg.Graph(64, eg.MODE_ADJ_SORT, undirected=True) as g:
        g.tune_set("wide_part", 10)
        print("wide empty graph", flush=True)
        g.load_edges(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
        ptr, col, _ = g.csr()
        assert g.info.m == 0 and not ptr.any() and col.size == 0
        dist, par, st = g.bfs(5)
        assert st.levels == 1 and st.reached == 1 and dist[5] == 1 and dist.sum() == 1
        assert par[5] == 5 and (np.delete(par, 5) == 64).all()


@pytest.mark.timeout(900)
def test_full_size_rmat27_undirected_symmetric():
    """R-MAT 27, edge factor 16, -u -s: 2^32 stored entries in two parts."""
    scale = 27
    n, m_in = 1 << scale, 16 << scale
    t0 = time.time()

    def step(msg):
        print(f"[{time.time() - t0:6.1f}s] {msg}", flush=True)

    step("full size: oracle records")
    s0, d0 = O.gen_rmat(scale, m_in, seed=1)
    with eg.Graph(n, eg.MODE_ADJ_SORT, undirected=True, symmetric=True) as g:
        step("device records + 64-bit-offset build")
        g.generate(eg.rmat_params(scale, seed=1))
        info = g.info
        assert info.m == 1 << 32 and info.has_csr == 1 and info.has_csc == 0
        step(f"built in {info.build_ms:.0f} ms, {info.device_bytes / 2**30:.1f} GiB")
        out_deg, in_deg = g.degrees()
        cnt, sm, xr = O.record_row_sums(n, s0, d0, 1)
        assert np.array_equal(out_deg, cnt) and np.array_equal(in_deg, cnt)
        step("degrees equal; downloading the CSR")
        ptr, col, _ = g.csr()
        assert ptr[0] == 0 and ptr[-1] == 1 << 32
        assert np.array_equal(np.diff(ptr), cnt)
        csum, cxr = O.csr_row_sums(n, ptr, col)
        assert np.array_equal(csum, sm) and np.array_equal(cxr, xr)
        step("row sums and xors equal")
        rng = np.random.default_rng(7)
        nz = np.nonzero(cnt)[0]
        sel = np.unique(np.concatenate([np.argsort(cnt)[-8:], rng.choice(nz, 56)])).astype(
            np.uint32)
        rp, rc = O.rows_of(n, s0, d0, 1, sel)
        for j, v in enumerate(sel):
            assert np.array_equal(col[ptr[v]:ptr[v + 1]], rc[rp[j]:rp[j + 1]]), int(v)
        step(f"{sel.size} sampled rows equal ({int(rp[-1])} entries)")
        hub = int(np.argmax(cnt))
        first = None
        for root, start_pull in ((hub, False), (int(rng.choice(nz)), False), (hub, True)):
            dist, par, st = g.bfs(root, start_pull=start_pull)
            step(f"BFS root {root} pull-first={start_pull}: {st.levels} levels "
                 f"({st.pull_levels} pull), {st.device_ms:.1f} ms")
            assert O.check_bfs_levels(n, s0, d0, 1, root, dist) == 0
            assert O.check_bfs_parents(n, ptr, col, root, dist, par) == 0
            assert st.levels == int(dist.max()) and st.reached == int((dist > 0).sum())
            if root == hub and first is None:
                first = dist
                assert st.pull_levels > 0
            elif root == hub:
                assert np.array_equal(dist, first)
        with pytest.raises(eg.EgError) as e:
            g.bfs_batch([hub])
        assert e.value.code == eg.EG_ENOTSUP
        step("done")
