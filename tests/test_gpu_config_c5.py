"""Parity at the full size of the serving configuration, against the oracle:
directed R-MAT 26 (67 M vertices, 1.07 G entries), CSR + CSC resident.  A
Poisson stream of 30 queries (80% BFS, 20% personalized PageRank with 1-8
seeds, tolerance 1e-6 L1) runs through the serving engine alone and in the
mix.  BFS results alone and in the mix are identical, four against the oracle;
PPR alone and in the mix agree within 2e-6 L1 (the load-adaptive engine may
take either kernel), and one query, alone and in the mix, lies within 1e-6 L1
of the FP64 oracle at the iteration count it ran.  Degrees of the device
layout equal the oracle's."""
import time

import numpy as np
import pytest

import eg_oracle_py as O
import everythinggraph_amd as eg

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[config parity] {msg}", flush=True)


# ---------------------------------------------------------------------------
# C5: directed R-MAT 26, serving a Poisson stream alone and in the mix
# ---------------------------------------------------------------------------
C5_SCALE = 26


@pytest.fixture(scope="module")
def c5():
    n = 1 << C5_SCALE
    say("C5: directed R-MAT 26, CSR + CSC, generated and built on the device")
    g = eg.Graph(n, eg.MODE_ADJ_SORT)
    g.generate(eg.rmat_params(C5_SCALE, 16, seed=1))
    assert g.info.m == 16 << C5_SCALE
    godeg, gideg = g.degrees()
    rng = np.random.default_rng(4)
    cand = np.flatnonzero(godeg).astype(np.uint32)
    qs = []
    for _ in range(30):
        if rng.random() < 0.8:
            qs.append(("bfs", int(rng.choice(cand))))
        else:
            qs.append(("ppr", rng.choice(cand, size=int(rng.integers(1, 9))).astype(np.uint32)))
    gaps = rng.exponential(1.0 / 20.0, len(qs))  # 20 queries/s offered

    def submit(S, q):
        return S.submit_bfs(q[1]) if q[0] == "bfs" else S.submit_ppr(q[1])

    say("C5: 30 queries alone")
    with g.serve() as S:
        alone = [S.wait(submit(S, q)) for q in qs]
    say("C5: the same 30 queries as a Poisson stream at 20 queries/s")
    with g.serve() as S:
        tickets = []
        t0 = time.perf_counter()
        due = 0.0
        for q, gap in zip(qs, gaps):
            due += gap
            wait = due - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            tickets.append(submit(S, q))
        mix = [S.wait(t) for t in tickets]
        st = S.stats()
    g.close()
    yield n, qs, alone, mix, st, godeg, gideg


@pytest.fixture(scope="module")
def c5_oracle(c5):
    n = c5[0]
    say("C5: oracle records, CSR and CSC")
    s, d = O.gen_rmat(C5_SCALE, 16 << C5_SCALE, seed=1)
    ptr, col, _ = O.build_csr(n, s, d)
    del s, d
    cp, row, _ = O.build_csc(n, ptr, col)
    yield ptr, col, cp, row


@pytest.mark.timeout(600)
def test_c5_alone_equals_mix(c5):
    n, qs, alone, mix, st, _, _ = c5
    assert st["completed"] == len(qs) and st["failed"] == 0
    for i, (q, (ia, oa), (ib, ob)) in enumerate(zip(qs, alone, mix)):
        if q[0] == "bfs":
            assert np.array_equal(oa, ob), i
            assert ia["levels"] == ib["levels"] and ia["reached"] == ib["reached"], i
        else:
            assert np.abs(oa.astype(np.float64) - ob).sum() <= 2e-6, i
            assert ia["delta"] <= 1e-6 and ib["delta"] <= 1e-6, i


@pytest.mark.timeout(600)
def test_c5_degrees(c5, c5_oracle):
    n, qs, alone, mix, st, godeg, gideg = c5
    ptr, col, cp, row = c5_oracle
    assert np.array_equal(godeg, np.diff(ptr).astype(np.uint32))
    assert np.array_equal(gideg, np.diff(cp).astype(np.uint32))


@pytest.mark.timeout(600)
def test_c5_bfs_vs_oracle(c5, c5_oracle):
    n, qs, alone, mix, st, _, _ = c5
    ptr, col, cp, row = c5_oracle
    bfs_i = [i for i, q in enumerate(qs) if q[0] == "bfs"][:4]
    assert len(bfs_i) == 4
    for i in bfs_i:
        say(f"C5: oracle BFS root {qs[i][1]}")
        ref, _, _ = O.bfs(n, int(ptr[-1]), ptr, col, cp, row, qs[i][1])
        assert np.array_equal(alone[i][1], ref), i
        assert alone[i][0]["reached"] == np.count_nonzero(ref), i


@pytest.mark.timeout(900)
def test_c5_ppr_vs_oracle(c5, c5_oracle):
    """The FP64 oracle takes about a second per iteration at this size: one
    query, the one that ran the fewest iterations, alone and in the mix."""
    n, qs, alone, mix, st, _, _ = c5
    ptr, col, cp, row = c5_oracle
    ppr_i = [i for i, q in enumerate(qs) if q[0] == "ppr"]
    assert ppr_i
    i = min(ppr_i, key=lambda j: alone[j][0]["levels"])
    deg = np.diff(ptr).astype(np.uint32)
    refs = {}
    for info, out in (alone[i], mix[i]):
        it = info["levels"]
        if it not in refs:
            say(f"C5: oracle PPR, {len(qs[i][1])} seeds, {it} iterations")
            refs[it] = O.ppr(n, cp, row, deg, qs[i][1], 0.15, it)
        assert np.abs(out.astype(np.float64) - refs[it]).sum() <= 1e-6, (i, it)
