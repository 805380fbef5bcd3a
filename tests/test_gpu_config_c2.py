"""Parity at the full size of the plain drop-in configuration, against the
oracle: R-MAT 22, edge factor 16, -u, CSR + CSC built on the device by the
count sort of load mode 4 (init_adj_count, init_all.c:444-496).  The layout
bit for bit; BFS levels from 64 seeded roots bit for bit with legal parent
trees (bfs_simple.c:136-229, the switch rule :191-197); WCC labels bit for bit
(wcc.c:30-42)."""
import numpy as np
import pytest

import eg_oracle_py as O
import everythinggraph_amd as eg

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[config parity] {msg}", flush=True)


# ---------------------------------------------------------------------------
# C2: R-MAT 22 -u, count-sort CSR, BFS from 64 roots, WCC
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2():
    scale = 22
    n = 1 << scale
    say("C2: R-MAT 22 -u, load mode 4, generated and built on the device")
    g = eg.Graph(n, eg.MODE_ADJ_COUNT, undirected=True)
    g.generate(eg.rmat_params(scale, 16, seed=1))
    say("C2: oracle records and layout")
    s0, d0 = O.gen_rmat(scale, 16 << scale, seed=1)
    s, d, _ = O.symmetrize(s0, d0)
    del s0, d0
    ptr, col, _ = O.build_csr(n, s, d)
    del s, d
    cp, row, _ = O.build_csc(n, ptr, col)
    yield g, n, ptr, col, cp, row
    g.close()


@pytest.mark.timeout(600)
def test_c2_count_sort_layout(c2):
    g, n, ptr, col, cp, row = c2
    gptr, gcol, _ = g.csr()
    assert np.array_equal(gptr, ptr) and np.array_equal(gcol, col)
    del gptr, gcol
    gcp, grow, _ = g.csc()
    assert np.array_equal(gcp, cp) and np.array_equal(grow, row)


@pytest.mark.timeout(600)
def test_c2_bfs_64_roots(c2):
    g, n, ptr, col, cp, row = c2
    m = int(ptr[-1])
    roots = eg.sample_roots(np.diff(ptr).astype(np.uint32), 64, seed=2)
    for i, root in enumerate(roots):
        if i % 16 == 0:
            say(f"C2: BFS roots {i}..{i + 15}")
        dist, par, st = g.bfs(int(root))
        ref, lv, _ = O.bfs(n, m, ptr, col, cp, row, int(root))
        assert np.array_equal(dist, ref), (i, int(root))
        assert st.levels == lv, i
        assert O.check_bfs_parents(n, ptr, col, int(root), dist, par) == 0, i


@pytest.mark.timeout(600)
def test_c2_wcc(c2):
    g, n, ptr, col, cp, row = c2
    say("C2: WCC")
    lab, _ = g.wcc()
    ref, _ = O.wcc_push(n, ptr, col)
    assert np.array_equal(lab, ref)
