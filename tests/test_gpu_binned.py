"""Binned (two-phase) PageRank on the device vs the FP64 oracle.

The binned kernels sum every row exactly in 64-bit fixed point, so their
ranks must match the oracle to the PageRank tolerance (L1 <= 1e-6, max error
<= 1e-6 of the largest rank), repeat bit for bit, and not depend on how rows
are cut into bins, sources into chunks, or vertices ordered, nor on the layout
the entries come from (CSR, 2-D grid, edge array): those only change the
order of integer additions.  SpMV sums the FP32 products w * x exactly and
is checked against the FP64 oracle at |y - y64| <= 1e-5 (|A| |x|).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import eg_oracle_py as O
import everythinggraph_amd as eg

pytestmark = pytest.mark.gpu

PR_L1_TOL = 1e-6
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rmat(scale, ef=16, seed=1):
    return O.gen_rmat(scale, ef << scale, seed=seed)


def oracle_pr(n, s, d, undirected, iters, dangling="drop", init=0.0):
    if undirected:
        s, d, _ = O.symmetrize(s, d)
    ptr, col, _ = O.build_csr(n, s, d)
    cp, row, _ = O.build_csc(n, ptr, col)
    deg = np.diff(ptr).astype(np.uint32)
    kw = {"dangling_uniform": dangling == "uniform"}
    if init:
        kw["init"] = init
    return O.pagerank_pull(n, cp, row, deg, 0.85, iters, **kw)


@pytest.mark.parametrize("scale,undirected", [(10, False), (12, True), (16, False), (17, True)])
@pytest.mark.parametrize("dangling", ["drop", "uniform"])
def test_binned_pagerank_matches_oracle(scale, undirected, dangling):
    n = 1 << scale
    s0, d0 = rmat(scale, seed=scale)
    ref, rdh, rlh = oracle_pr(n, s0, d0, undirected, 20, dangling)
    with eg.Graph(n, 2, undirected=undirected) as g:
        g.load_edges(s0, d0)
        g.tune_set("pr_binned", 1)
        r, dh, lh = g.pagerank(0.85, 20, dangling=dangling)
        r2, _, _ = g.pagerank(0.85, 20, dangling=dangling)
        geo = g.binned_geometry()
        g.tune_set("pr_binned", 0)
        rp, _, _ = g.pagerank(0.85, 20, dangling=dangling)
    err = np.abs(r.astype(np.float64) - ref)
    assert err.sum() <= PR_L1_TOL, err.sum()
    assert err.max() <= 1e-6 * ref.max()
    assert np.array_equal(r, r2)  # deterministic
    np.testing.assert_allclose(dh, rdh, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(lh, rlh, rtol=1e-3, atol=1e-7)
    # the pull kernel agrees to FP32 rounding; the exact sums are no further
    # from the FP64 oracle than the compensated ones
    assert np.abs(r.astype(np.float64) - rp).sum() <= PR_L1_TOL
    assert err.sum() <= np.abs(rp.astype(np.float64) - ref).sum() * 1.5 + 1e-9
    # hot sources (read from LDS in phase 2) come in windows of 24320; the
    # rest are scattered in chunks
    assert geo["bins"] >= 1 and geo["items"] >= 1  # an item groups bins or is a long row's part
    assert geo["hot"] == n or geo["hot"] % 24320 == 0
    assert (geo["chunks"] >= 1) == (geo["hot"] < n)


def test_binned_reference_defaults():
    """prev = 0.15 per vertex and 10 iterations (pagerank_simple.c:95,115):
    the rank mass starts at 0.15 n, and the fixed-point scale must hold it."""
    n = 1 << 12
    s0, d0 = rmat(12, seed=7)
    ref, _, _ = oracle_pr(n, s0, d0, False, 10, init=0.15)
    with eg.Graph(n, 2) as g:
        g.load_edges(s0, d0)
        g.tune_set("pr_binned", 1)
        r, _, _ = g.pagerank(0.85, 10, init=0.15)
    np.testing.assert_allclose(r, ref, rtol=2e-6, atol=1e-9)


def test_binned_closed_forms():
    """Directed cycle: every rank stays 1/n.  Star into a hub (a row with
    n - 1 entries, summed in parts): the hub's rank is (1-d)/n + d (n-1)/n *
    r_leaf with leaves at (1-d)/n after the first iteration."""
    n = 4096
    src = np.arange(n, dtype=np.uint32)
    dst = ((src + 1) % n).astype(np.uint32)
    with eg.Graph(n, 2) as g:
        g.load_edges(src, dst)
        g.tune_set("pr_binned", 1)
        r, _, _ = g.pagerank(0.85, 20)
    np.testing.assert_allclose(r, 1.0 / n, rtol=1e-6)

    n = 1 << 15
    leaves = np.arange(1, n, dtype=np.uint32)
    with eg.Graph(n, 2) as g:
        g.load_edges(leaves, np.zeros(n - 1, np.uint32))
        g.tune_set("pr_binned", 1)
        r, _, _ = g.pagerank(0.85, 5)
    base = 0.15 / n
    np.testing.assert_allclose(r[1:], base, rtol=1e-6)
    np.testing.assert_allclose(r[0], base + 0.85 * (n - 1) * base, rtol=1e-6)


@pytest.mark.parametrize("undirected", [False, True])
def test_binned_single_vector_ppr(undirected):
    """The plain per-query PPR on the binned kernels: within 1e-6 L1 of the
    FP64 oracle at the same iteration count, and of the pull kernel."""
    n = 1 << 13
    s0, d0 = rmat(13, seed=41)
    s, d = (O.symmetrize(s0, d0)[:2]) if undirected else (s0, d0)
    ptr, col, _ = O.build_csr(n, s, d)
    cp, row, _ = O.build_csc(n, ptr, col)
    deg = np.diff(ptr).astype(np.uint32)
    with eg.Graph(n, 2, undirected=undirected) as g:
        g.load_edges(s0, d0)
        out = {}
        for pb in (1, 0):
            g.tune_set("pr_binned", pb)
            out[pb] = [g.ppr(sd, 0.15, tol=0.0, max_iter=30) for sd in ([3], [1, 2, 2, 900])]
    for q, sd in enumerate(([3], [1, 2, 2, 900])):
        ref = O.ppr(n, cp, row, deg, sd, 0.15, 30)
        r, it, _ = out[1][q]
        assert it == 30
        assert np.abs(r - ref).sum() <= PR_L1_TOL
        assert np.abs(r.astype(np.float64) - out[0][q][0]).sum() <= PR_L1_TOL


_GEOMETRY_SCRIPT = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[1] + '/tests')
import eg_oracle_py as O, everythinggraph_amd as eg
n = 1 << 13
s, d = O.gen_rmat(13, 16 << 13, seed=31)
out = {}
for und in (False, True):
    with eg.Graph(n, 2, undirected=und) as g:
        g.load_edges(s, d)
        g.tune_set("pr_binned", 1)
        out['pr%d' % und] = g.pagerank(0.85, 40)[0]
        out['uni%d' % und] = g.pagerank(0.85, 40, dangling="uniform")[0]
        geo = g.binned_geometry()
        out['geo%d' % und] = np.array([geo[k] for k in ("bins", "chunks", "tiles", "items",
                                                       "long_rows")])
        g.tune_set("pr_hot", 0)      # every source scattered
        out['cold%d' % und] = g.pagerank(0.85, 40)[0]
        g.tune_set("pr_hot", 4096)   # half hot, half scattered
        out['half%d' % und] = g.pagerank(0.85, 40)[0]
        g.tune_set("pr_xcd", 1)      # scatter tiles dealt to XCDs in contiguous ranges
        out['xcd%d' % und] = g.pagerank(0.85, 40)[0]
        g.tune_set("pr_persist", 0)  # a phase-2 workgroup per work item
        out['item%d' % und] = g.pagerank(0.85, 40)[0]
np.savez(sys.argv[2], **out)
"""


def test_binned_independent_of_cut_and_order(tmp_path):
    """The same ranks, bit for bit, with the default bins and chunks, with
    tiny ones (40 hot sources in three windows of 16, long rows in parts,
    tiles cut by length and by bins), with all, half or none of the sources
    hot, and with or without the degree-ordered copy (built for runs of 32
    iterations or more), in either scatter tile order, and with persistent
    phase-2 workgroups or one per work item."""
    script = tmp_path / "run.py"
    script.write_text(_GEOMETRY_SCRIPT)
    res = {}
    for geo in ("default", "small"):
        for order in ("none", "degree"):
            env = dict(os.environ, EG_PULL_ORDER=order, EG_PB_GEOMETRY=geo)
            outp = tmp_path / f"{geo}_{order}.npz"
            print(f"binned geometry={geo} order={order}", flush=True)
            subprocess.run([sys.executable, str(script), ROOT, str(outp)], env=env, check=True,
                           timeout=120)
            res[(geo, order)] = np.load(outp)
    first = res[("default", "none")]
    for key, r in res.items():
        for k in ("pr0", "pr1", "uni0", "uni1"):
            assert np.array_equal(r[k], first[k]), (key, k)
        for und in (0, 1):  # no hot sources, or half of them: the same bits
            assert np.array_equal(r[f"cold{und}"], first[f"pr{und}"]), (key, und)
            assert np.array_equal(r[f"half{und}"], first[f"pr{und}"]), (key, und)
            assert np.array_equal(r[f"xcd{und}"], first[f"pr{und}"]), (key, und)
            assert np.array_equal(r[f"item{und}"], first[f"pr{und}"]), (key, und)
    for und in (0, 1):
        small = res[("small", "degree")][f"geo{und}"]
        bins, chunks, tiles, items, long_rows = (int(v) for v in small)
        assert long_rows > 0 and items != bins  # hub rows in parts, short bins grouped
        assert tiles > chunks                  # chunk regions cut into several tiles
        assert bins > 16                       # tiles also cut by the bins they span


LAYOUTS = [(2, 0), (0, 16), (1, 4), (6, 25), (7, 0), (8, 0), (9, 0)]


@pytest.mark.parametrize("undirected", [False, True])
def test_binned_pagerank_every_layout(undirected):
    """Adjacency, grid (modes 0/1/6) and edge-array (7/8/9) PageRank on the
    binned kernels: within the tolerance of the oracle and bit-identical to
    each other (hot and scattered sources both present at this size)."""
    n = 1 << 15
    s0, d0 = rmat(15, seed=61)
    ref, rdh, rlh = oracle_pr(n, s0, d0, undirected, 20)
    out = {}
    for mode, Pp in LAYOUTS:
        with eg.Graph(n, mode, undirected=undirected, P=Pp) as g:
            g.load_edges(s0, d0)
            g.tune_set("pr_binned", 1)
            out[mode] = g.pagerank(0.85, 20)
    for mode, (r, dh, lh) in out.items():
        err = np.abs(r.astype(np.float64) - ref)
        assert err.sum() <= PR_L1_TOL and err.max() <= 1e-6 * ref.max(), mode
        assert np.array_equal(r, out[2][0]), mode
        np.testing.assert_allclose(dh, rdh, rtol=1e-5, atol=1e-12)
        np.testing.assert_allclose(lh, rlh, rtol=1e-3, atol=1e-7)


def spmv_check(y, ref, absref):
    assert np.all(np.abs(y.astype(np.float64) - ref) <= 1e-5 * absref + 1e-30)


@pytest.mark.parametrize("weighted", [True, False])
def test_binned_spmv_every_layout(weighted):
    """SpMV y = A^T x with signed x on the binned kernels, every layout:
    within 1e-5 (|A| |x|) of the FP64 oracle and bit-identical across
    layouts (the same FP32 products, summed exactly)."""
    n = 1 << 15
    s0, d0 = rmat(15, seed=62)
    w = O.gen_weights(s0.size, seed=3) if weighted else None
    ptr, col, ow = O.build_csr(n, s0, d0, w)
    cp, row, iw = O.build_csc(n, ptr, col, ow)
    x = np.random.default_rng(7).uniform(-2, 2, n).astype(np.float32)
    x[::97] = 0.0
    ref = O.spmv_pull(n, cp, row, iw, x.astype(np.float64))
    absref = O.spmv_pull(n, cp, row, iw, np.abs(x).astype(np.float64))
    out = {}
    for mode, Pp in LAYOUTS:
        with eg.Graph(n, mode, weighted=weighted, P=Pp) as g:
            g.load_edges(s0, d0, w)
            g.tune_set("pr_binned", 1)
            out[mode] = g.spmv(x)
            if mode == 2:
                geo = g.binned_geometry()
    for mode, y in out.items():
        spmv_check(y, ref, absref)
        assert np.array_equal(y, out[2]), mode
    assert geo["hot"] == n  # the first sources of an R-MAT graph take many entries


@pytest.mark.parametrize("weighted", [False, True])
def test_binned_hot_windows(weighted):
    """Hot sources in several LDS windows of 24320 (R-MAT 16: 65536 sources):
    one window (partial), two (the last one partial) and three, each with one phase-2
    workgroup per work item and with persistent workgroups (which walk the
    windows in alternating directions), give the PageRank ranks and the SpMV
    products of the single-window plan bit for bit, and the ranks stay within
    the oracle's tolerance."""
    scale = 16
    n = 1 << scale
    s0, d0 = rmat(scale, seed=71)
    w = O.gen_weights(s0.size, seed=5) if weighted else None
    ref, _, _ = oracle_pr(n, s0, d0, False, 20)
    x = np.random.default_rng(9).uniform(-2, 2, n).astype(np.float32)
    out = {}
    with eg.Graph(n, 2, weighted=weighted) as g:
        g.load_edges(s0, d0, w)
        g.tune_set("pr_binned", 1)
        for hot in (16384, 40000, 65536):
            for persist in (0, 1):
                print(f"hot windows: weighted={weighted} pr_hot={hot} persist={persist}",
                      flush=True)
                g.tune_set("pr_hot", hot)
                g.tune_set("pr_persist", persist)
                r, _, _ = g.pagerank(0.85, 20)
                y = g.spmv(x)
                assert g.binned_geometry()["hot"] == hot
                out[(hot, persist)] = (r, y)
    r0, y0 = out[(16384, 0)]
    err = np.abs(r0.astype(np.float64) - ref)
    assert err.sum() <= PR_L1_TOL and err.max() <= 1e-6 * ref.max()
    for key, (r, y) in out.items():
        assert np.array_equal(r, r0), key
        assert np.array_equal(y, y0), key


def test_binned_spmv_nonfinite_falls_back():
    """An x with inf / NaN has no finite fixed-point scale: the pull kernel
    runs instead and propagates them as IEEE arithmetic does."""
    n = 1 << 12
    s0, d0 = rmat(12, seed=63)
    x = np.ones(n, np.float32)
    x[5] = np.inf
    x[9] = np.nan
    with eg.Graph(n, 2) as g:
        g.load_edges(s0, d0)
        g.tune_set("pr_binned", 0)
        y0 = g.spmv(x)
        g.tune_set("pr_binned", 1)
        y1 = g.spmv(x)
    assert np.array_equal(y0, y1, equal_nan=True)
    assert not np.isfinite(y1).all()


@pytest.mark.parametrize("mode", [2, 0])
@pytest.mark.parametrize("case", ["inf_weight", "nan_weight", "product_overflow"])
def test_binned_spmv_nonfinite_weights_fall_back(mode, case):
    """A weight of inf or NaN (12-byte records may hold them), or a finite
    weight whose FP32 product with x overflows, has no finite fixed-point
    scale either: the pull / grid kernel runs, on the adjacency lists and on
    the grid, with the same IEEE results as pr_binned = 0."""
    n = 1 << 12
    s0, d0 = rmat(12, seed=65)
    w = O.gen_weights(s0.size, seed=3)
    x = np.random.default_rng(3).uniform(-1, 1, n).astype(np.float32)
    if case == "inf_weight":
        w[7] = np.inf
    elif case == "nan_weight":
        w[11] = np.nan
    else:
        w[13] = 3e38
        x[:] = 2.0  # 6e38 > FLT_MAX: the product is inf in FP32
    with eg.Graph(n, mode, weighted=True, P=16 if mode == 0 else 0) as g:
        g.load_edges(s0, d0, w)
        layout = "grid" if mode == 0 else "adj"
        g.tune_set("pr_binned", 0)
        y0 = g.spmv(x, layout=layout)
        g.tune_set("pr_binned", 1)
        y1 = g.spmv(x, layout=layout)
        # finite weights again: the binned kernels sum exactly
        w2 = np.where(np.isfinite(w) & (w < 1.0), w, 0.5).astype(np.float32)
        g.load_edges(s0, d0, w2)
        x2 = np.random.default_rng(4).uniform(-1, 1, n).astype(np.float32)
        ya = g.spmv(x2, layout=layout)
        g.tune_set("pr_binned", 0)
        yb = g.spmv(x2, layout=layout)
    assert np.array_equal(y0, y1, equal_nan=True)
    assert not np.isfinite(y1).all()
    ptr, col, ow = O.build_csr(n, s0, d0, w2)
    cp, row, iw = O.build_csc(n, ptr, col, ow)
    ref = O.spmv_pull(n, cp, row, iw, x2.astype(np.float64))
    absref = O.spmv_pull(n, cp, row, iw, np.abs(x2).astype(np.float64))
    for y in (ya, yb):
        assert np.all(np.abs(y.astype(np.float64) - ref) <= 1e-5 * absref + 1e-30)


def test_binned_uniform_graph_has_no_hot_set():
    """Erdos-Renyi ids carry no degree skew: every window of 24320 sources
    takes 1.2% of the entries (under the 2% a window must hold to be kept
    beyond the first), and the first alone is under 15%, so the plan scatters
    every source."""
    n = 1 << 21
    s0, d0 = O.gen_er(n, 1 << 22, seed=64)
    ptr, col, _ = O.build_csr(n, s0, d0)
    cp, row, _ = O.build_csc(n, ptr, col)
    ref, _, _ = O.pagerank_pull(n, cp, row, np.diff(ptr).astype(np.uint32), 0.85, 10)
    with eg.Graph(n, 0, P=16) as g:
        g.load_edges(s0, d0)
        g.tune_set("pr_binned", 1)
        r, _, _ = g.pagerank(0.85, 10)
        geo = g.binned_geometry()
    assert geo["hot"] == 0 and geo["chunks"] > 1
    assert np.abs(r.astype(np.float64) - ref).sum() <= PR_L1_TOL


def test_batched_ppr_static_tail():
    """Batched PPR on a symmetric graph whose degree order ends in isolated
    vertices (rows without entries or out-edges): after two full steps the
    steps skip those rows.  Seeds include an isolated vertex; results stay
    within 1e-6 L1 of the oracle for iteration counts on both sides of the
    switch, eager and replayed from a hipGraph alike."""
    n = 1 << 13
    s0, d0 = rmat(13, seed=65)
    s, d, _ = O.symmetrize(s0, d0)
    ptr, col, _ = O.build_csr(n, s, d)
    cp, row, _ = O.build_csc(n, ptr, col)
    deg = np.diff(ptr).astype(np.uint32)
    iso = int(np.flatnonzero(deg == 0)[0])
    seeds = [[iso], [1, iso], [7], [100, 3, 3, 9]]
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        for graphs in (0, 1):
            g.tune_set("graphs", graphs)
            for it in (1, 2, 3, 6, 7, 30):
                out = g.ppr_batch(seeds, 0.15, it)[0]
                for q, sd in enumerate(seeds):
                    ref = O.ppr(n, cp, row, deg, sd, 0.15, it)
                    assert np.abs(out[q] - ref).sum() <= PR_L1_TOL, (graphs, it, q)


def test_auto_choice_weighs_preparation():
    """pr_binned = -1 builds the binned plan (and the degree-ordered copy)
    only for runs long enough to repay it: the reference's 10 iterations
    (pagerank_simple.c:115) on a graph of 2^26 entries build neither and give
    the pull kernel's bits; a 200-iteration run builds both, and from then on
    the plan is used."""
    scale = 21
    n = 1 << scale
    with eg.Graph(n, eg.MODE_ADJ_SORT, undirected=True, symmetric=True) as g:
        g.generate(eg.rmat_params(scale, 16, seed=1))
        assert g.info.m >= 1 << 22
        r10, _, _ = g.pagerank(0.85, 10)
        p = g.prep
        assert p["binned_plan_ms"] == 0 and p["order_ms"] == 0
        g.tune_set("pr_binned", 0)
        assert np.array_equal(r10, g.pagerank(0.85, 10)[0])
        g.tune_set("pr_binned", -1)
        g.pagerank(0.85, 200, download=False)
        p = g.prep
        assert p["binned_plan_ms"] > 0 and p["order_ms"] > 0
        assert p["binned_plan_bytes"] > 8 * g.info.m
        r = g.pagerank(0.85, 10)[0]
        g.tune_set("pr_binned", 1)
        assert np.array_equal(r, g.pagerank(0.85, 10)[0])


def test_drop_in_pagerank_takes_the_pull_path(tmp_path):
    """The drop-in pagerank_simple (load + the reference's 10 iterations from
    0.15) on a graph past the binned size threshold runs the pull kernel: its
    ranks equal pr_binned = 0's bit for bit."""
    scale = 21
    n = 1 << scale
    s, d = O.gen_rmat(scale, 16 << scale, seed=2)
    rec = np.empty((s.size, 2), np.uint32)
    rec[:, 0], rec[:, 1] = s, d
    path = tmp_path / "rmat21.bin"
    path.write_bytes(rec.tobytes())
    del rec
    out = tmp_path / "rank.out"
    exe = os.path.join(eg.BIN_DIR, "pagerank_simple")
    print("driver: pagerank_simple on R-MAT 21 -u", flush=True)
    r = subprocess.run([exe, "-f", str(path), "-n", str(n), "-m", "2", "-u"], capture_output=True,
                       text=True, timeout=300, env={**os.environ, "EG_DUMP": str(out)})
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got = np.fromfile(out, np.float32)
    with eg.Graph(n, eg.MODE_ADJ_SORT, undirected=True) as g:
        g.load_edges(s, d)
        g.tune_set("pr_binned", 0)
        ref, _, _ = g.pagerank(0.85, 10, init=0.15)
    assert np.array_equal(got, ref)
