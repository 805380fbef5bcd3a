"""Algorithms on the device vs the oracle.

Integer results (BFS levels, WCC labels) are bit-exact; BFS parents are
checked as a legal BFS tree.  PageRank, SpMV and PPR run in FP32 with
compensated sums and are compared with the FP64 oracle:
  * PageRank: L1 error <= 1e-6 and max error <= 1e-6 relative to max rank,
    for ranks initialised to 1/n (sum ~ 1);
  * SpMV: |y - y64| <= 1e-5 * (|A| |x|) elementwise;
  * PPR: L1 error <= 1e-6 per query vector (each sums to <= 1).
"""
import os

import numpy as np
import pytest

import eg_oracle_py as O
import everythinggraph_amd as eg

pytestmark = pytest.mark.gpu

PR_L1_TOL = 1e-6


def rmat(scale, ef=16, seed=1):
    return O.gen_rmat(scale, ef << scale, seed=seed)


def oracle_adj(n, s, d, w=None, undirected=False):
    if undirected:
        s, d, w = O.symmetrize(s, d, w)
    ptr, col, ow = O.build_csr(n, s, d, w)
    cp, row, iw = O.build_csc(n, ptr, col, ow)
    return s, d, w, ptr, col, ow, cp, row, iw


# ---------------------------------------------------------------------------
# PageRank
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("scale,undirected", [(10, False), (12, True), (16, False)])
@pytest.mark.parametrize("dangling", ["drop", "uniform"])
def test_pagerank_adjacency(scale, undirected, dangling):
    n = 1 << scale
    s0, d0 = rmat(scale)
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=undirected)
    deg = np.diff(ptr).astype(np.uint32)
    ref, rdh, rlh = O.pagerank_pull(n, cp, row, deg, 0.85, 20, dangling_uniform=dangling == "uniform")
    with eg.Graph(n, 2, undirected=undirected) as g:
        g.load_edges(s0, d0)
        r, dh, lh = g.pagerank(0.85, 20, dangling=dangling)
        r2, _, _ = g.pagerank(0.85, 20, dangling=dangling)
    err = np.abs(r.astype(np.float64) - ref)
    assert err.sum() <= PR_L1_TOL, err.sum()
    assert err.max() <= 1e-6 * ref.max()
    assert np.array_equal(r, r2)  # deterministic
    np.testing.assert_allclose(dh, rdh, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(lh, rlh, rtol=1e-3, atol=1e-7)


def test_pagerank_config_c1():
    """pagerank_simple on RMAT-16 EF16 directed, FP64 oracle, r0 = 1/n, 20 it."""
    scale = 16
    n = 1 << scale
    s0, d0 = rmat(scale)
    _, _, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0)
    ref, _, _ = O.pagerank_pull(n, cp, row, np.diff(ptr).astype(np.uint32), 0.85, 20)
    with eg.Graph(n, 2) as g:
        g.generate(eg.rmat_params(scale, 16, seed=1))
        r, _, _ = g.pagerank(0.85, 20)
    err = np.abs(r.astype(np.float64) - ref)
    assert err.sum() <= PR_L1_TOL and err.max() <= 1e-6 * ref.max()


def test_pagerank_reference_defaults():
    """The drop-in defaults: prev = 0.15, 10 iterations (pagerank_simple.c:95,115)."""
    n = 1 << 10
    s0, d0 = rmat(10)
    _, _, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0)
    ref, _, _ = O.pagerank_pull(n, cp, row, np.diff(ptr).astype(np.uint32), 0.85, 10, init=0.15)
    with eg.Graph(n, 2) as g:
        g.load_edges(s0, d0)
        r, _, _ = g.pagerank(0.85, 10, init=0.15)
    np.testing.assert_allclose(r, ref, rtol=2e-6, atol=1e-9)


def test_pagerank_symmetric_uses_csr():
    n = 1 << 11
    s0, d0 = rmat(11)
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=True)
    ref, _, _ = O.pagerank_pull(n, cp, row, np.diff(ptr).astype(np.uint32), 0.85, 20)
    with eg.Graph(n, 2, undirected=True, symmetric=True) as g:
        g.load_edges(s0, d0)
        r, _, _ = g.pagerank(0.85, 20)
    assert np.abs(r - ref).sum() <= PR_L1_TOL


@pytest.mark.parametrize("mode,Pp", [(0, 16), (1, 4), (6, 25)])
def test_pagerank_grid(mode, Pp):
    n = 1 << 12
    s0, d0 = rmat(12)
    _, _, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0)
    ref, _, rlh = O.pagerank_pull(n, cp, row, np.diff(ptr).astype(np.uint32), 0.85, 20)
    with eg.Graph(n, mode, P=Pp) as g:
        g.load_edges(s0, d0)
        r, _, lh = g.pagerank(0.85, 20)
    assert np.abs(r - ref).sum() <= PR_L1_TOL
    np.testing.assert_allclose(lh, rlh, rtol=1e-3, atol=1e-7)


@pytest.mark.parametrize("mode", [7, 8, 9])
def test_pagerank_edges(mode):
    n = 1 << 12
    s0, d0 = rmat(12)
    _, _, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0)
    ref, _, _ = O.pagerank_pull(n, cp, row, np.diff(ptr).astype(np.uint32), 0.85, 20)
    with eg.Graph(n, mode) as g:
        g.load_edges(s0, d0)
        for binned in (0, 1):  # FP64-accumulated edge push, exact binned sums
            g.tune_set("pr_binned", binned)
            r, _, _ = g.pagerank(0.85, 20)
            err = np.abs(r.astype(np.float64) - ref)
            assert err.sum() <= PR_L1_TOL and err.max() <= 1e-6 * ref.max(), binned


def test_pagerank_closed_form_cycle():
    n = 1000
    s = np.arange(n, dtype=np.uint32)
    d = ((s + 1) % n).astype(np.uint32)
    with eg.Graph(n, 2) as g:
        g.load_edges(s, d)
        r, _, lh = g.pagerank(0.85, 10)
    np.testing.assert_allclose(r, 1.0 / n, rtol=1e-6)


# ---------------------------------------------------------------------------
# SpMV
# ---------------------------------------------------------------------------
def spmv_check(y, ref, absref):
    assert np.all(np.abs(y.astype(np.float64) - ref) <= 1e-5 * absref + 1e-30)


@pytest.mark.parametrize("mode,Pp", [(2, 0), (0, 16), (6, 7), (1, 3), (8, 0), (9, 0)])
def test_spmv_weighted(mode, Pp):
    n = 1 << 12
    s0, d0 = rmat(12, seed=4)
    w = O.gen_weights(s0.size, seed=2)
    _, _, _, ptr, col, ow, cp, row, iw = oracle_adj(n, s0, d0, w)
    x = np.random.default_rng(1).uniform(-1, 1, n).astype(np.float32)
    ref = O.spmv_pull(n, cp, row, iw, x.astype(np.float64))
    absref = O.spmv_pull(n, cp, row, iw, np.abs(x).astype(np.float64))
    with eg.Graph(n, mode, weighted=True, P=Pp) as g:
        g.load_edges(s0, d0, w)
        y = g.spmv(x)
    spmv_check(y, ref, absref)


def test_spmv_unweighted_and_hub():
    n = 3000
    rng = np.random.default_rng(5)
    s = rng.integers(0, n, 200000).astype(np.uint32)
    d = rng.integers(0, n, 200000).astype(np.uint32)
    d[:90000] = 7  # a long in-row split across segments
    _, _, _, ptr, col, _, cp, row, _ = oracle_adj(n, s, d)
    x = rng.uniform(0, 1, n).astype(np.float32)
    ref = O.spmv_pull(n, cp, row, None, x.astype(np.float64))
    with eg.Graph(n, 2) as g:
        g.load_edges(s, d)
        y = g.spmv(x)
    spmv_check(y, ref, ref)


# ---------------------------------------------------------------------------
# BFS
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("scale,undirected", [(10, False), (12, True), (14, True), (14, False)])
def test_bfs_adjacency(scale, undirected):
    n = 1 << scale
    s0, d0 = rmat(scale, seed=2)
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=undirected)
    deg = np.diff(ptr)
    roots = eg.sample_roots(deg, 6, seed=2)
    with eg.Graph(n, 2, undirected=undirected) as g:
        g.load_edges(s0, d0)
        for root in roots:
            ref, lv, _ = O.bfs(n, s.size, ptr, col, cp, row, int(root))
            for start_pull in (False, True):
                for switch in (False, True):
                    dist, par, st = g.bfs(int(root), start_pull, switch)
                    assert np.array_equal(dist, ref)
                    assert st.levels == lv
                    assert O.check_bfs_parents(n, ptr, col, int(root), dist, par) == 0
            reached = ref != 0
            assert st.reached == reached.sum()
            esum = int(deg[reached].sum())
            assert st.edges_traversed == (esum // 2 if undirected else esum)


def test_bfs_pull_path_is_taken():
    n = 1 << 14
    s0, d0 = rmat(14, seed=3)
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=True)
    root = int(eg.sample_roots(np.diff(ptr), 1, seed=9)[0])
    ref, _, _ = O.bfs(n, s.size, ptr, col, cp, row, root)
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        dist, par, st = g.bfs(root, False, True)
    assert st.pull_levels > 0 and np.array_equal(dist, ref)
    assert O.check_bfs_parents(n, ptr, col, root, dist, par) == 0


def test_bfs_long_diameter():
    """A 3000-vertex chain into a dense cluster: thousands of levels (many
    batches of device-side levels), then pull levels at the far end."""
    rng = np.random.default_rng(12)
    chain = np.arange(3000, dtype=np.uint32)
    cs = rng.integers(3000, 6000, 40000).astype(np.uint32)
    cd = rng.integers(3000, 6000, 40000).astype(np.uint32)
    s0 = np.concatenate([chain, cs])
    d0 = np.concatenate([chain + 1, cd])
    n = 6000
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=True)
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        for root in (0, 1500, 4000):
            ref, lv, _ = O.bfs(n, s.size, ptr, col, cp, row, root)
            for start_pull, switch in ((False, True), (True, True), (True, False)):
                dist, par, st = g.bfs(root, start_pull, switch)
                assert np.array_equal(dist, ref) and st.levels == lv, (root, start_pull, switch)
                assert O.check_bfs_parents(n, ptr, col, root, dist, par) == 0
            assert lv > 1000 or root == 4000
        dist, par, st = g.bfs(0, False, True)
        assert st.pull_levels > 0  # the cluster's frontier crosses m/20


def test_bfs_symmetric_and_zero_degree():
    # directed graph whose frontier holds sinks; also -s (no CSC) on -u input
    n = 2000
    s = np.array([0, 0, 0, 1, 2, 5], np.uint32)
    d = np.array([1, 2, 3, 4, 5, 6], np.uint32)
    _, _, _, ptr, col, _, cp, row, _ = oracle_adj(n, s, d)
    ref, lv, _ = O.bfs(n, s.size, ptr, col, cp, row, 0)
    with eg.Graph(n, 2) as g:
        g.load_edges(s, d)
        for sp_ in (False, True):
            dist, par, st = g.bfs(0, sp_, True)
            assert np.array_equal(dist, ref) and st.levels == lv
    s0, d0 = rmat(12, seed=5)
    us, ud, _, ptr, col, _, cp, row, _ = oracle_adj(1 << 12, s0, d0, undirected=True)
    with eg.Graph(1 << 12, 2, undirected=True, symmetric=True) as g:
        g.load_edges(s0, d0)
        root = int(us[0])
        ref, _, _ = O.bfs(1 << 12, us.size, ptr, col, ptr, col, root)
        dist, par, _ = g.bfs(root, True, True)
        assert np.array_equal(dist, ref)


@pytest.mark.parametrize("mode", [0, 1, 6, 7, 8, 9])
def test_bfs_grid_and_edges(mode):
    n = 1 << 12
    s0, d0 = rmat(12, seed=6)
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=True)
    roots = eg.sample_roots(np.diff(ptr), 3, seed=1)
    with eg.Graph(n, mode, undirected=True, P=16) as g:
        g.load_edges(s0, d0)
        for root in roots:
            ref, lv, _ = O.bfs(n, s.size, ptr, col, cp, row, int(root))
            dist, par, st = g.bfs(int(root))
            assert np.array_equal(dist, ref) and st.levels == lv
            assert O.check_bfs_parents(n, ptr, col, int(root), dist, par) == 0


# ---------------------------------------------------------------------------
# WCC
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("undirected", [False, True])
@pytest.mark.parametrize("scale", [10, 14])
def test_wcc_adjacency(undirected, scale):
    n = 1 << scale
    s0, d0 = rmat(scale, ef=2, seed=7)
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=undirected)
    ref, _ = O.wcc_push(n, ptr, col)
    with eg.Graph(n, 4, undirected=undirected) as g:
        g.load_edges(s0, d0)
        lab, rounds = g.wcc()
    assert np.array_equal(lab, ref)


@pytest.mark.parametrize("symmetric", [False, True])
def test_wcc_ordered_pull_equals_edges(symmetric):
    """Undirected graphs: pull rounds on the degree-ordered ids (wcc_order 1)
    and edge relaxation in the layout's ids (0) give the oracle's labels, the
    minimum caller id of each component, with many small components beside
    the one holding the top hub (rows that never see label 0) and hub rows
    longer than one lane's scan."""
    n = 1 << 14
    s0, d0 = rmat(14, ef=2, seed=9)
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=True)
    ref, _ = O.wcc_push(n, ptr, col)
    assert np.diff(ptr).max() > 64 and np.unique(ref).size > 100
    with eg.Graph(n, 2, undirected=True, symmetric=symmetric) as g:
        g.load_edges(s0, d0)
        for order in (1, 0, 1):
            g.tune_set("wcc_order", order)
            lab, _ = g.wcc()
            assert np.array_equal(lab, ref), order


@pytest.mark.parametrize("mode", [0, 1, 6, 7, 8, 9])
def test_wcc_grid_and_edges(mode):
    n = 1 << 12
    s0, d0 = rmat(12, ef=1, seed=8)
    ref, _ = O.wcc_edges(n, s0, d0)
    with eg.Graph(n, mode, P=9) as g:
        g.load_edges(s0, d0)
        lab, _ = g.wcc()
    assert np.array_equal(lab, ref)


# ---------------------------------------------------------------------------
# batched form
# ---------------------------------------------------------------------------
def test_bfs_batch_equals_single():
    n = 1 << 13
    s0, d0 = rmat(13, seed=11)
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=True)
    roots = eg.sample_roots(np.diff(ptr), 70, seed=2)  # two groups (64 + 6)
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        out, ms = g.bfs_batch(roots)
        for i, r in enumerate(roots):
            ref, _, _ = O.bfs(n, s.size, ptr, col, cp, row, int(r))
            assert np.array_equal(out[i], ref), i


@pytest.mark.parametrize("undirected,symmetric,k", [(True, True, 200), (False, False, 130)])
def test_bfs_batch_two_lane_words(undirected, symmetric, k):
    """More than 64 sources run in traversals of 128 lanes (two mask words per
    vertex); push levels (symmetric graphs) and pull levels both carry the
    second word, and a last group smaller than 64 leaves it idle."""
    n = 1 << 13
    s0, d0 = rmat(13, seed=14)
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=undirected)
    roots = eg.sample_roots(np.diff(ptr), k, seed=4)
    with eg.Graph(n, 2, undirected=undirected, symmetric=symmetric) as g:
        g.load_edges(s0, d0)
        out, _ = g.bfs_batch(roots)
    for i, r in enumerate(roots):
        ref, _, _ = O.bfs(n, s.size, ptr, col, cp, row, int(r))
        assert np.array_equal(out[i], ref), i


def test_bfs_batch_directed_duplicates():
    n = 1 << 11
    s0, d0 = rmat(11, seed=12)
    _, _, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0)
    roots = np.array([0, 0, 5, int(s0[3])], np.uint32)
    with eg.Graph(n, 2) as g:
        g.load_edges(s0, d0)
        out, _ = g.bfs_batch(roots)
    for i, r in enumerate(roots):
        ref, _, _ = O.bfs(n, s0.size, ptr, col, cp, row, int(r))
        assert np.array_equal(out[i], ref)


def test_ppr_batch_vs_oracle():
    n = 1 << 12
    s0, d0 = rmat(12, seed=13)
    _, _, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=True)
    deg = np.diff(ptr).astype(np.uint32)
    rng = np.random.default_rng(3)
    seeds = [rng.choice(n, size=rng.integers(1, 9), replace=True).astype(np.uint32)
             for _ in range(66)]
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        out, ms = g.ppr_batch(seeds, 0.15, 30)
    for q, sd in enumerate(seeds):
        ref = O.ppr(n, cp, row, deg, sd, 0.15, 30)
        assert np.abs(out[q] - ref).sum() <= PR_L1_TOL, q


# ---------------------------------------------------------------------------
# degree-ordered compute copy: results identical with and without it
# ---------------------------------------------------------------------------
_ORDER_SCRIPT = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[1] + '/tests')
import eg_oracle_py as O, everythinggraph_amd as eg
n = 1 << 13
s, d = O.gen_rmat(13, 16 << 13, seed=21)
w = O.gen_weights(s.size, seed=3)
x = np.linspace(-1, 1, n).astype(np.float32)
out = {}
with eg.Graph(n, 2, weighted=True) as g:
    g.load_edges(s, d, w)
    out['pr'] = g.pagerank(0.85, 40)[0]  # long enough to build the degree-ordered copy
    out['spmv'] = g.spmv(x)
with eg.Graph(n, 2, undirected=True) as g:
    g.load_edges(s, d)
    out['ppr'] = g.ppr_batch([[1, 2], [5]], 0.15, 30)[0]
    out['ms'] = g.bfs_batch([0, 7, 99])[0]
np.savez(sys.argv[2], **out)
"""


def test_degree_order_is_bit_identical(tmp_path):
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = tmp_path / "run.py"
    script.write_text(_ORDER_SCRIPT)
    res = {}
    for mode in ("none", "degree"):
        env = dict(os.environ, EG_PULL_ORDER=mode)
        outp = tmp_path / f"{mode}.npz"
        subprocess.run([sys.executable, str(script), root, str(outp)], env=env, check=True,
                       timeout=120)
        res[mode] = np.load(outp)
    for k in ("pr", "spmv", "ppr", "ms"):
        assert np.array_equal(res["none"][k], res["degree"][k]), k


# ---------------------------------------------------------------------------
# kernel shapes: every variant gives bit-identical results
# ---------------------------------------------------------------------------
def test_tuning_variants_bit_identical():
    n = 1 << 13
    s0, d0 = rmat(13, seed=51)
    _, _, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=True)
    roots = eg.sample_roots(np.diff(ptr), 40, seed=2)
    seeds = [[1, 2], [7], [100, 3, 3, 9]]
    outs = {}
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        for v in (8, 16):
            for p in ("pull_batch", "spmm_depth", "ms_batch"):
                g.tune_set(p, v)
                assert g.tune_get(p) == v
            g.tune_set("pull_hot", 0 if v == 8 else n // 16)
            g.tune_set("pull_persist", 0 if v == 8 else 2)
            g.tune_set("ms_push_div", 20 if v == 8 else 2)  # more levels pushed
            outs[v] = (g.pagerank(0.85, 20)[0], g.ppr_batch(seeds, 0.15, 20)[0],
                       g.bfs_batch(roots)[0], g.ppr([5, 6], 0.15, tol=0.0, max_iter=25)[0])
        with pytest.raises(eg.EgError) as e:
            g.tune_set("pull_batch", 12)
        assert e.value.code == eg.EG_EINVAL
        with pytest.raises(eg.EgError) as e:
            g.tune_set("no_such_knob", 8)
        assert e.value.code == eg.EG_EINVAL
        with pytest.raises(eg.EgError) as e:
            g.tune_set("ms_push_div", 0)
        assert e.value.code == eg.EG_EINVAL
    for a, b in zip(outs[8], outs[16]):
        assert np.array_equal(a, b)
    with eg.Graph(n, 2, undirected=True) as g:  # one persistent workgroup per CU
        g.load_edges(s0, d0)
        g.tune_set("pull_persist", 1)
        assert np.array_equal(g.pagerank(0.85, 20)[0], outs[8][0])
        assert np.array_equal(g.ppr([5, 6], 0.15, tol=0.0, max_iter=25)[0], outs[8][3])
    for i, r in enumerate(roots):
        ref, _, _ = O.bfs(n, s0.size * 2, ptr, col, cp, row, int(r))
        assert np.array_equal(outs[16][2][i], ref)


def test_autotune_reports_and_keeps_fastest():
    n = 1 << 12
    s0, d0 = rmat(12, seed=52)
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        rep = g.autotune()
        params = {r[0] for r in rep}
        # the hot-set and persistent pull shapes exist only on the degree-ordered copy
        ordered = os.environ.get("EG_PULL_ORDER") != "none"
        expect = {"pull_batch", "pr_binned", "pr_xcd", "pr_hot", "pr_persist", "spmm_depth",
                  "ms_batch", "ms_push_div"}
        if ordered:
            expect |= {"pull_hot", "pull_persist"}
        assert params == expect
        assert sorted(r[1] for r in rep if r[0] == "pr_binned") == [0, 1]
        assert sorted(r[1] for r in rep if r[0] == "pr_xcd") == [0, 1]
        assert sorted(r[1] for r in rep if r[0] == "ms_push_div") == [5, 20, 40]
        # 4096 sources fit one hot window: every larger cap gives the same plan
        assert [r[1] for r in rep if r[0] == "pr_hot"] == [24320]
        assert sorted(r[1] for r in rep if r[0] == "pr_persist") == [0, 1]
        if ordered:
            hot = sorted(r[1] for r in rep if r[0] == "pull_hot")
            assert hot == [0, n // 64, n // 16, n // 4]
        for p in sorted(expect):
            rows = [r for r in rep if r[0] == p]
            assert all(r[2] > 0 for r in rows)
            if p in ("pull_batch", "spmm_depth", "ms_batch"):
                assert sorted(r[1] for r in rows) == [8, 16]
            best = min(rows, key=lambda r: r[2])
            chosen = [r for r in rows if r[3]]
            assert len(chosen) == 1 and chosen[0][1] == g.tune_get(p) == best[1]
        r, _, _ = g.pagerank(0.85, 20)  # still runs after tuning
        assert np.isfinite(r).all()


def test_ppr_graph_replay_matches_eager():
    """Batched PPR iterations replayed from a hipGraph equal eager launches,
    on every replay, with other device work between the calls."""
    n = 1 << 12
    s0, d0 = rmat(12, seed=53)
    _, _, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=True)
    deg = np.diff(ptr).astype(np.uint32)
    seeds = [[1, 2], [7], [100, 3, 3, 9], [4000]]
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        g.tune_set("graphs", 0)
        eager = {it: g.ppr_batch(seeds, 0.15, it)[0] for it in (4, 6, 7, 30)}
        g.tune_set("graphs", 1)
        for rep in range(3):
            for it in (4, 6, 7, 30):
                out = g.ppr_batch(seeds, 0.15, it)[0]
                assert np.array_equal(out, eager[it]), (rep, it)
            g.pagerank(0.85, 3)          # unrelated eager work between replays
            g.bfs(1)
    for q, sd in enumerate(seeds):
        ref = O.ppr(n, cp, row, deg, sd, 0.15, 30)
        assert np.abs(eager[30][q] - ref).sum() <= PR_L1_TOL


def test_ppr_batch_without_static_rows():
    """Batched PPR steps before the last write only the contribution rows;
    on a graph where every vertex has edges (a directed cycle through all
    vertices plus R-MAT records: no static rows to skip) the last step alone
    must produce every row, eager and replayed alike."""
    n = 1 << 12
    s0, d0 = rmat(12, seed=57)
    ring = np.arange(n, dtype=np.uint32)
    s0 = np.concatenate([s0, ring])
    d0 = np.concatenate([d0, (ring + 1) % n]).astype(np.uint32)
    _, _, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=False)
    deg = np.diff(ptr).astype(np.uint32)
    assert deg.min() >= 1
    seeds = [[1, 2], [7], [100, 3, 3, 9], [4000]]
    with eg.Graph(n, 2) as g:
        g.load_edges(s0, d0)
        for graphs in (0, 1):
            g.tune_set("graphs", graphs)
            for it in (3, 4, 6, 7, 30):
                out = g.ppr_batch(seeds, 0.15, it)[0]
                for q, sd in enumerate(seeds):
                    ref = O.ppr(n, cp, row, deg, sd, 0.15, it)
                    assert np.abs(out[q] - ref).sum() <= PR_L1_TOL, (graphs, it, q)


@pytest.mark.parametrize("mode", [0, 6])
def test_grid_bands_bit_identical(mode):
    """Grid PageRank / SpMV give the same bits for every source-band width."""
    n = 1 << 12
    s0, d0 = rmat(12, seed=54)
    w = O.gen_weights(s0.size, seed=5)
    x = np.linspace(-1, 1, n).astype(np.float32)
    outs = []
    with eg.Graph(n, mode, weighted=True, P=7) as g:
        g.load_edges(s0, d0, w)
        for band in (1, 2, 3, 0):
            g.tune_set("grid_band", band)
            outs.append((g.pagerank(0.85, 20, layout="grid")[0], g.spmv(x, layout="grid")))
        rep = g.autotune(eg.TUNE_GRID)
        bands = [r for r in rep if r[0] == "grid_band"]
        assert sorted(r[1] for r in bands) == [0, 1, 2, 4] and sum(r[3] for r in bands) == 1
        assert sorted(r[1] for r in rep if r[0] == "pr_binned") == [0, 1]
    for pr, y in outs[1:]:
        assert np.array_equal(pr, outs[0][0]) and np.array_equal(y, outs[0][1])
    _, _, _, ptr, col, ow, cp, row, iw = oracle_adj(n, s0, d0, w)
    deg = np.diff(ptr).astype(np.uint32)
    ref, _, _ = O.pagerank_pull(n, cp, row, deg, 0.85, 20)
    assert np.abs(outs[0][0] - ref).sum() <= PR_L1_TOL


# ---------------------------------------------------------------------------
# SSSP (sssp_pushpull.c)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("scale,undirected,mode", [(10, False, 2), (12, True, 2), (11, False, 4),
                                                   (11, True, 3)])
@pytest.mark.parametrize("weights", ["ref", "unit"])
def test_sssp_vs_oracle(scale, undirected, mode, weights):
    n = 1 << scale
    s0, d0 = rmat(scale, seed=60 + scale)
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=undirected)
    kind = eg.SSSP_WEIGHTS[weights]
    roots = eg.sample_roots(np.diff(ptr), 4, seed=2)
    with eg.Graph(n, mode, undirected=undirected) as g:
        g.load_edges(s0, d0)
        for r in roots:
            ref, _ = O.sssp(n, ptr, col, int(r), kind)
            for start_pull, switch in ((False, True), (True, True), (False, False)):
                dist, par, st = g.sssp(int(r), start_pull, switch, weights)
                assert np.array_equal(dist, ref), (int(r), start_pull, switch)
                assert O.check_sssp_parents(n, ptr, col, int(r), kind, dist, par) == 0
                assert st.reached == np.count_nonzero(ref != O.SSSP_UNREACHED)
                if start_pull and switch:
                    assert st.pull_iterations >= 1
            if weights == "unit":  # unit weights are BFS levels minus one
                lv, _, _ = O.bfs(n, s.size, ptr, col, cp, row, int(r))
                assert np.array_equal(np.where(lv == 0, O.SSSP_UNREACHED, lv - 1), ref)


@pytest.mark.parametrize("weights", ["ref", "unit"])
def test_sssp_hub_rows(weights):
    """In-rows long enough to be cut into segments (9000 and 3000 entries) and
    rows reduced by a whole wave (500 entries), on every direction schedule."""
    n = 4096
    rng = np.random.default_rng(61)
    s = rng.integers(0, n, 80000).astype(np.uint32)
    d = rng.integers(0, n, 80000).astype(np.uint32)
    d[:9000], d[9000:12000], d[12000:12500] = 7, 100, 200
    _, _, _, ptr, col, _, _, _, _ = oracle_adj(n, s, d)
    kind = eg.SSSP_WEIGHTS[weights]
    with eg.Graph(n, 2) as g:
        g.load_edges(s, d)
        for r in (int(s[0]), int(s[9500]), 7):
            ref, _ = O.sssp(n, ptr, col, r, kind)
            for start_pull, switch in ((True, False), (True, True), (False, True)):
                dist, par, st = g.sssp(r, start_pull, switch, weights)
                assert np.array_equal(dist, ref), (r, start_pull, switch)
                assert O.check_sssp_parents(n, ptr, col, r, kind, dist, par) == 0


def test_sssp_symmetric_flag_and_errors():
    n = 1 << 10
    s0, d0 = rmat(10, seed=66)
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=True)
    with eg.Graph(n, 2, undirected=True, symmetric=True) as g:  # CSR only, pulls over it
        g.load_edges(s0, d0)
        ref, _ = O.sssp(n, ptr, col, 3, 0)
        assert np.array_equal(g.sssp(3, start_pull=True)[0], ref)
        with pytest.raises(eg.EgError) as e:
            g.sssp(n)
        assert e.value.code == eg.EG_ERANGE
    with eg.Graph(n, 0) as g:
        g.load_edges(s0, d0)
        with pytest.raises(eg.EgError) as e:
            g.sssp(0)
        assert e.value.code == eg.EG_ENOTSUP
    with eg.Graph(n, 2) as g:  # isolated root and an empty graph
        g.load_edges(np.array([5], np.uint32), np.array([6], np.uint32))
        dist, par, st = g.sssp(9)
        assert st.reached == 1 and dist[9] == 0 and par[9] == 9
        assert np.all(np.delete(dist, 9) == O.SSSP_UNREACHED)


@pytest.mark.parametrize("symmetric", [False, True])
def test_bfs_order_same_levels(symmetric):
    """Symmetric graphs: BFS over rows ordered hubs first (bfs_order 1), over
    the degree-ordered ids with rows hubs first (2) and in input order (0)
    give the same levels, in the caller's ids, and legal parent trees."""
    n = 1 << 13
    s0, d0 = rmat(13, seed=66)
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=True)
    roots = eg.sample_roots(np.diff(ptr), 5, seed=3)
    with eg.Graph(n, 2, undirected=True, symmetric=symmetric) as g:
        g.load_edges(s0, d0)
        for root in roots:
            out = {}
            for order in (0, 1, 2):
                g.tune_set("bfs_order", order)
                out[order] = g.bfs(int(root))
            ref, _, st = O.bfs(n, s.size, ptr, col, cp, row, int(root))
            for order in (0, 1, 2):
                assert np.array_equal(out[order][0], ref), order
                assert out[order][2].reached == out[0][2].reached
                assert out[order][2].edges_traversed == out[0][2].edges_traversed
                assert O.check_bfs_parents(n, ptr, col, int(root), out[order][0],
                                           out[order][1]) == 0


@pytest.mark.parametrize("symmetric", [False, True])
def test_bfs_ordered_ids_isolated_roots(symmetric):
    """Ordered ids: vertices without edges are the tail of the relabelled ids,
    which pull levels do not scan.  Traversals from such roots (alone in
    their component) and from ordinary roots, pulled from the first level or
    not, in any order, give the oracle's levels."""
    n = 1 << 13
    s0, d0 = rmat(13, seed=67)
    s, d, _, ptr, col, _, cp, row, _ = oracle_adj(n, s0, d0, undirected=True)
    deg = np.diff(ptr)
    iso = np.nonzero(deg == 0)[0]
    assert iso.size > 2
    roots = eg.sample_roots(deg, 3, seed=5)
    seq = [(int(iso[0]), False), (int(roots[0]), False), (int(iso[-1]), True),
           (int(roots[1]), True), (int(iso[1]), False), (int(roots[2]), False)]
    with eg.Graph(n, 2, undirected=True, symmetric=symmetric) as g:
        g.load_edges(s0, d0)
        g.tune_set("bfs_order", 2)
        for r, start_pull in seq:
            dist, par, st = g.bfs(r, start_pull=start_pull)
            ref, _, _ = O.bfs(n, s.size, ptr, col, cp, row, r)
            assert np.array_equal(dist, ref), (r, start_pull)
            assert O.check_bfs_parents(n, ptr, col, r, dist, par) == 0
            if deg[r] == 0:
                assert st.reached == 1 and dist[r] == 1
