"""Parity at the full size of the 2-D grid configuration, against the oracle:
Erdos-Renyi 2^24 vertices / 2^28 directed edges with FP32 weights in (0, 1],
16 x 16 grid (init_grid_sort, init_all.c:252-361).  The grid bit for bit; 20
PageRank iterations (prgrid_cilk.c:62-76) on the binned kernels and on the
grid-tile kernel with L1 <= 1e-6 and max <= 1e-6 max(r) against FP64; SpMV
(spmv.c:98-106 with weights) with |y - y64| <= 1e-5 (|A| |x|) per row."""
import numpy as np
import pytest

import eg_oracle_py as O
import everythinggraph_amd as eg

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[config parity] {msg}", flush=True)


# ---------------------------------------------------------------------------
# C3: ER 2^24 / 2^28 weighted, 16 x 16 grid, PageRank and SpMV
# ---------------------------------------------------------------------------
C3_N, C3_M, C3_P = 1 << 24, 1 << 28, 16


@pytest.fixture(scope="module")
def c3():
    n, m, Pp = C3_N, C3_M, C3_P
    say("C3: ER 2^24 / 2^28 weighted, 16 x 16 grid, built on the device")
    g = eg.Graph(n, eg.MODE_GRID_SORT, weighted=True, P=Pp)
    g.generate(eg.er_params(m, seed=1, weight_seed=1))
    say("C3: oracle records and grid")
    s, d = O.gen_er(n, m, seed=1)
    w = O.gen_weights(m, seed=1)
    cp, gs, gd, gw = O.build_grid(n, Pp, s, d, w)
    odeg = O.out_degree(n, s)
    del s, d, w
    yield g, cp, gs, gd, gw, odeg
    g.close()


@pytest.mark.timeout(600)
def test_c3_grid_layout(c3):
    g, cp, gs, gd, gw, odeg = c3
    assert g.info.P == C3_P
    gcp, ggs, ggd, ggw = g.grid()
    assert np.array_equal(gcp, cp) and np.array_equal(ggs, gs)
    assert np.array_equal(ggd, gd) and np.array_equal(ggw, gw)
    del gcp, ggs, ggd, ggw
    gdeg, _ = g.degrees()
    assert np.array_equal(gdeg, odeg)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("binned", [1, 0])  # the binned kernels, the grid-tile kernel
def test_c3_grid_pagerank(c3, binned):
    g, cp, gs, gd, gw, odeg = c3
    say(f"C3: FP64 oracle PageRank, 20 iterations; pr_binned={binned}")
    ref = O.pagerank_grid(C3_N, C3_P, cp, gs, gd, odeg, 0.85, 20)
    g.tune_set("pr_binned", binned)
    r, _, _ = g.pagerank(0.85, 20, layout="grid")
    err = np.abs(r.astype(np.float64) - ref)
    assert err.sum() <= 1e-6, err.sum()
    assert err.max() <= 1e-6 * ref.max(), err.max()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("binned", [1, 0])
def test_c3_grid_spmv(c3, binned):
    g, cp, gs, gd, gw, odeg = c3
    say(f"C3: FP64 oracle SpMV; pr_binned={binned}")
    x = np.random.default_rng(7).uniform(-1, 1, C3_N).astype(np.float32)
    yref = O.spmv_grid(C3_N, C3_P, cp, gs, gd, gw, x)
    yabs = O.spmv_grid(C3_N, C3_P, cp, gs, gd, gw, np.abs(x))
    g.tune_set("pr_binned", binned)
    y = g.spmv(x, layout="grid")
    assert np.all(np.abs(y.astype(np.float64) - yref) <= 1e-5 * yabs + 1e-30)
