"""Serving form (eg_serve_*): continuous batching must not change any result.

Every query is run three ways: alone (submitted only after the previous one
finished), in a mix submitted at once, and in a mix whose arrivals are
staggered so that queries join traversals and SpMM iterations already in
flight.  BFS levels must equal the oracle's bit for bit in every run.  With
the engine always batching (idle_plain = -1) a PPR vector must be
bit-identical across the runs, take the same number of iterations, and lie
within L1 1e-6 of the FP64 oracle run for that many iterations (the vector
sums to <= 1).  With the default load-adaptive engine a query that finds it
idle runs on the single-vector kernel: its PPR vector is held to the same
oracle bound, and agrees with the batched one within 2e-6 L1.
"""
import time

import numpy as np
import pytest

import eg_oracle_py as O
import everythinggraph_amd as eg

pytestmark = pytest.mark.gpu

PPR_L1_TOL = 1e-6


def build(scale, seed, undirected):
    n = 1 << scale
    s0, d0 = O.gen_rmat(scale, 16 << scale, seed=seed)
    s, d = (O.symmetrize(s0, d0)[:2] if undirected else (s0, d0))
    ptr, col, _ = O.build_csr(n, s, d)
    cp, row, _ = O.build_csc(n, ptr, col)
    return n, s0, d0, s.size, ptr, col, cp, row


def queries(n, deg, count, seed=4, bfs_fraction=0.8):
    rng = np.random.default_rng(seed)
    cand = np.flatnonzero(deg).astype(np.uint32)
    out = []
    for _ in range(count):
        if rng.random() < bfs_fraction:
            out.append(("bfs", int(rng.choice(cand))))
        else:
            out.append(("ppr", rng.choice(cand, size=int(rng.integers(1, 9))).astype(np.uint32)))
    return out


def submit(S, q):
    kind, arg = q
    return S.submit_bfs(arg) if kind == "bfs" else S.submit_ppr(arg)


def run_alone(g, qs, **opts):
    with g.serve(**opts) as S:
        return [S.wait(submit(S, q)) for q in qs]


def run_mix(g, qs, stagger_s=0.0, **opts):
    with g.serve(**opts) as S:
        tickets = []
        for i, q in enumerate(qs):
            tickets.append(submit(S, q))
            if stagger_s and i % 5 == 4:
                time.sleep(stagger_s)
        res = [S.wait(t) for t in tickets]
        st = S.stats()
    return res, st


def check_against_oracle(n, m, ptr, col, cp, row, qs, res):
    deg = np.diff(ptr).astype(np.uint32)
    for i, (q, (info, out)) in enumerate(zip(qs, res)):
        if q[0] == "bfs":
            ref, _, _ = O.bfs(n, m, ptr, col, cp, row, q[1])
            assert np.array_equal(out, ref), (i, q[1])
            assert info["reached"] == np.count_nonzero(ref), i
            assert info["levels"] == ref.max(), i
        else:
            ref = O.ppr(n, cp, row, deg, q[1], 0.15, info["levels"])
            assert np.abs(out.astype(np.float64) - ref).sum() <= PPR_L1_TOL, i
            assert info["delta"] <= 1e-6 or info["levels"] == 200, i


def check_same(qs, a, b):
    for i, (q, (ia, oa), (ib, ob)) in enumerate(zip(qs, a, b)):
        assert np.array_equal(oa, ob), (i, q[0])
        assert ia["levels"] == ib["levels"], (i, q[0])
        if q[0] == "ppr":
            assert ia["delta"] == ib["delta"], i


@pytest.mark.parametrize("scale,undirected", [(12, True), (11, False)])
def test_alone_equals_mix(scale, undirected):
    n, s0, d0, m, ptr, col, cp, row = build(scale, 31 + scale, undirected)
    qs = queries(n, np.diff(ptr), 120)
    with eg.Graph(n, 2, undirected=undirected) as g:
        g.load_edges(s0, d0)
        alone = run_alone(g, qs, idle_plain=-1)
        mix, st = run_mix(g, qs, idle_plain=-1)
        stag, _ = run_mix(g, qs, stagger_s=0.002, idle_plain=-1)
        # three SpMM workgroups walk every work item (grid-stride)
        tiny, _ = run_mix(g, qs, stagger_s=0.001, ppr_grid=3, idle_plain=-1)
    check_against_oracle(n, m, ptr, col, cp, row, qs, alone)
    check_same(qs, alone, mix)
    check_same(qs, alone, stag)
    check_same(qs, alone, tiny)
    nb = sum(q[0] == "bfs" for q in qs)
    assert st["completed"] == len(qs) and st["failed"] == 0
    assert st["bfs_inflight"] == 0 and st["ppr_inflight"] == 0 and st["pending"] == 0
    # batching happened: fewer level sweeps than BFS queries x depth
    assert 0 < st["bfs_levels"] < nb * 3
    assert st["ppr_steps"] > 0 and st["device_bytes"] > 0
    assert st["served_plain"] == 0


def test_load_adaptive_default():
    """At light load a query runs on the plain kernels (BFS single-source, PPR
    single-vector); under a burst the engine batches.  Levels stay exact and
    PPR stays within the oracle bound either way."""
    n, s0, d0, m, ptr, col, cp, row = build(12, 39, True)
    qs = queries(n, np.diff(ptr), 60)
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        with g.serve() as S:
            alone = [S.wait(submit(S, q)) for q in qs]
            st_alone = S.stats()
        mix, st_mix = run_mix(g, qs)
        batched, _ = run_mix(g, qs, idle_plain=-1)
    assert st_alone["served_plain"] == len(qs)  # every lone query found it idle
    check_against_oracle(n, m, ptr, col, cp, row, qs, alone)
    check_against_oracle(n, m, ptr, col, cp, row, qs, mix)
    for i, (q, (ia, oa), (ib, ob)) in enumerate(zip(qs, alone, batched)):
        if q[0] == "bfs":
            assert np.array_equal(oa, ob), i
        else:
            assert np.abs(oa.astype(np.float64) - ob).sum() <= 2e-6, i


def test_many_lane_groups():
    n, s0, d0, m, ptr, col, cp, row = build(12, 41, True)
    roots = eg.sample_roots(np.diff(ptr), 150, seed=2)
    qs = [("bfs", int(r)) for r in roots]
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        res, st = run_mix(g, qs, bfs_lanes=192)
    assert st["bfs_lanes"] >= 128
    check_against_oracle(n, m, ptr, col, cp, row, qs, res)


def test_plain_mode_same_results():
    n, s0, d0, m, ptr, col, cp, row = build(11, 43, True)
    qs = queries(n, np.diff(ptr), 30)
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        plain, st = run_mix(g, qs, plain=True)
        mix, _ = run_mix(g, qs)
    # plain PPR runs the single-vector pull kernel (another summation order):
    # both are held to the oracle, BFS levels must agree exactly
    check_against_oracle(n, m, ptr, col, cp, row, qs, plain)
    check_against_oracle(n, m, ptr, col, cp, row, qs, mix)
    for i, (q, (ia, oa), (ib, ob)) in enumerate(zip(qs, plain, mix)):
        if q[0] == "bfs":
            assert np.array_equal(oa, ob), i
    assert st["ppr_columns"] == 0  # no SpMM block in plain mode


@pytest.mark.parametrize("undirected", [True, False])
def test_single_vector_ppr(undirected):
    n, s0, d0, m, ptr, col, cp, row = build(12, 48, undirected)
    deg = np.diff(ptr).astype(np.uint32)
    with eg.Graph(n, 2, undirected=undirected) as g:
        g.load_edges(s0, d0)
        for sd in ([3], [9, 9, 17, 250], list(range(40, 48))):
            out, it, delta = g.ppr(sd, 0.15, tol=1e-6, max_iter=200)
            ref = O.ppr(n, cp, row, deg, sd, 0.15, it)
            assert np.abs(out.astype(np.float64) - ref).sum() <= PPR_L1_TOL
            assert delta <= 1e-6 and 0 < it < 200
            # tol 0 runs exactly max_iter iterations
            out2, it2, _ = g.ppr(sd, 0.15, tol=0.0, max_iter=30)
            ref2 = O.ppr(n, cp, row, deg, sd, 0.15, 30)
            assert it2 == 30 and np.abs(out2.astype(np.float64) - ref2).sum() <= PPR_L1_TOL
        with pytest.raises(eg.EgError) as e:
            g.ppr([n])
        assert e.value.code == eg.EG_ERANGE


def test_iteration_cap_matches_batch():
    """A PPR query capped at k iterations equals eg_ppr_batch with k."""
    n, s0, d0, m, ptr, col, cp, row = build(11, 44, True)
    seeds = [[1, 2, 3], [7], [100, 100, 5]]
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        ref, _ = g.ppr_batch(seeds, 0.15, 7)
        with g.serve(ppr_tol=1e-300, ppr_max_iter=7, idle_plain=-1) as S:
            ts = [S.submit_ppr(s) for s in seeds]
            res = [S.wait(t) for t in ts]
    for q, (info, out) in enumerate(res):
        assert info["levels"] == 7
        assert np.array_equal(out, ref[q]), q


def test_errors_and_ownership():
    n, s0, d0, *_ = build(10, 45, True)
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        S = g.serve()
        try:
            with pytest.raises(eg.EgError) as e:
                g.bfs(0)
            assert e.value.code == eg.EG_ESTATE
            with pytest.raises(eg.EgError) as e:
                S.submit_bfs(n)
            assert e.value.code == eg.EG_ERANGE
            with pytest.raises(eg.EgError) as e:
                S.submit_ppr([])
            assert e.value.code == eg.EG_EINVAL
            with pytest.raises(eg.EgError) as e:
                S.wait(12345)
            assert e.value.code == eg.EG_EINVAL
            t = S.submit_bfs(0, download=False)
            info, out = S.wait(t)
            assert out is None and info["reached"] >= 1
        finally:
            S.stop()
        g.bfs(0)  # usable again
        with pytest.raises(eg.EgError) as e:
            g.serve(ppr_columns=65)
        assert e.value.code == eg.EG_EINVAL


def test_no_device_memory_for_ppr_fails_cleanly():
    """A reserve larger than the device leaves no room for any slot."""
    n, s0, d0, *_ = build(10, 46, True)
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        with g.serve(hbm_reserve=1 << 60, idle_plain=-1) as S:
            tb, tp = S.submit_bfs(0), S.submit_ppr([1])
            for t in (tb, tp):
                with pytest.raises(eg.EgError) as e:
                    S.wait(t)
                assert e.value.code == eg.EG_ENOMEM


@pytest.mark.parametrize("plain", [False, True])
def test_replay_report(plain):
    n, s0, d0, *_ = build(12, 47, True)
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        r = g.serve_replay(qps=4000.0, requests=300, plain=plain)
    assert r["completed"] == 300 and r["failed"] == 0
    assert r["bfs"] + r["ppr"] == 300 and r["bfs"] > r["ppr"] > 0
    assert 0 < r["p50_ms"] <= r["p99_ms"] <= r["max_ms"]
    assert r["qps"] > 0 and r["wall_s"] > 0
    assert r["stats"]["completed"] == 300


@pytest.mark.parametrize("binned", [0, 1])
@pytest.mark.parametrize("undirected", [True, False])
def test_solo_pipeline_equals_single_vector(binned, undirected):
    """A lone PPR query runs on the engine's single-vector path, which keeps a
    second iteration queued while the first is in flight and takes it back
    when the first ends the query.  Results must equal eg_ppr (one iteration
    at a time) bit for bit, with the same iteration count and last L1 change:
    tol 10 ends every query at iteration 1 with iteration 2 queued, a tiny
    tol runs to the iteration cap."""
    n, s0, d0, m, ptr, col, cp, row = build(11, 52, undirected)
    seeds = ([3], [9, 9, 17, 250], list(range(40, 48)))
    with eg.Graph(n, 2, undirected=undirected) as g:
        g.load_edges(s0, d0)
        g.tune_set("pr_binned", binned)
        for tol, cap in ((10.0, 200), (0.05, 200), (1e-3, 200), (1e-6, 200), (1e-300, 5)):
            ref = [g.ppr(sd, 0.15, tol=tol, max_iter=cap) for sd in seeds]
            with g.serve(ppr_tol=tol, ppr_max_iter=cap) as S:
                got = [S.wait(S.submit_ppr(sd)) for sd in seeds]
                st = S.stats()
            assert st["served_plain"] == len(seeds) and st["ppr_columns"] == 0
            for q, ((info, out), (rout, rit, rdelta)) in enumerate(zip(got, ref)):
                assert info["levels"] == rit, (tol, q)
                assert info["delta"] == rdelta, (tol, q)
                assert np.array_equal(out, rout), (tol, q)
            if tol == 10.0:
                assert all(info["levels"] == 1 for info, _ in got)
            if cap == 5:
                assert all(info["levels"] == 5 for info, _ in got)


def test_wait_while_submitting():
    """A waiter blocked on one ticket while another thread submits hundreds of
    queries (the ticket table grows and rehashes under it) gets its result,
    and every query matches the oracle."""
    import threading
    n, s0, d0, m, ptr, col, cp, row = build(11, 47, True)
    qs = queries(n, np.diff(ptr), 300, seed=9)
    k = next(i for i, q in enumerate(qs) if q[0] == "ppr")
    qs.insert(0, qs.pop(k))  # the waiter's query runs longest
    with eg.Graph(n, 2, undirected=True) as g:
        g.load_edges(s0, d0)
        with g.serve() as S:
            first = submit(S, qs[0])
            box = {}
            waiter = threading.Thread(target=lambda: box.setdefault("r", S.wait(first)))
            waiter.start()
            tickets = [submit(S, q) for q in qs[1:]]
            waiter.join(60)
            assert not waiter.is_alive()
            res = [box["r"]] + [S.wait(t) for t in tickets]
    check_against_oracle(n, m, ptr, col, cp, row, qs, res)
