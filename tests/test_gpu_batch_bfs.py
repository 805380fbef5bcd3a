"""Multi-source BFS (eg_bfs_batch) with device-driven levels.

The level kernels decide each level's direction by the reference's rule
(bfs_simple.c:191-197) on the frontier of all lanes and end the traversal
themselves; the host enqueues levels in batches (replayed from a hipGraph of
two levels) and waits once per batch.  The levels of every lane must equal
the oracle's queue BFS bit for bit, and the device-driven forms (graph
replay, eager launches) must equal the host-driven one (a host decision and
two host waits per level) for 64, 128 and 256 lanes per traversal, on
undirected graphs (push and pull levels) and directed ones (pull only), with
a last group that leaves lanes idle.  A traversal deeper than the u16 level
store (65534 levels) is rerun source by source, in the batch call and in the
serving engine.
"""
import numpy as np
import pytest

import eg_oracle_py as O
import everythinggraph_amd as eg

pytestmark = pytest.mark.gpu


def graph_and_oracle(scale, seed, undirected):
    n = 1 << scale
    s0, d0 = O.gen_rmat(scale, 16 << scale, seed=seed)
    s, d = (O.symmetrize(s0, d0)[:2] if undirected else (s0, d0))
    ptr, col, _ = O.build_csr(n, s, d)
    cp, row, _ = O.build_csc(n, ptr, col)
    return n, s0, d0, s.size, ptr, col, cp, row


@pytest.mark.parametrize("lanes", [64, 128, 256])
@pytest.mark.parametrize("undirected,symmetric", [(True, True), (True, False), (False, False)])
def test_device_levels_equal_host_levels_and_oracle(lanes, undirected, symmetric):
    n, s0, d0, m, ptr, col, cp, row = graph_and_oracle(12, 71 + lanes, undirected)
    roots = eg.sample_roots(np.diff(ptr).astype(np.uint32), 300, seed=5)  # a partial last group
    outs, infos = {}, {}
    with eg.Graph(n, eg.MODE_ADJ_SORT, undirected=undirected, symmetric=symmetric) as g:
        g.load_edges(s0, d0)
        g.tune_set("ms_lanes", lanes)
        for form, (dev, graphs) in {"graph": (1, 1), "eager": (1, 0), "host": (0, 0)}.items():
            g.tune_set("ms_device", dev)
            g.tune_set("graphs", graphs)
            print(f"batch BFS lanes={lanes} form={form}", flush=True)
            outs[form], _ = g.bfs_batch(roots)
            infos[form] = g.bfs_batch_info()
        # a second graph-replayed call reuses the recording (same lanes)
        g.tune_set("ms_device", 1)
        g.tune_set("graphs", 1)
        again, _ = g.bfs_batch(roots)
    assert np.array_equal(outs["graph"], outs["host"])
    assert np.array_equal(outs["eager"], outs["host"])
    assert np.array_equal(again, outs["host"])
    for i, r in enumerate(roots):
        ref, _, _ = O.bfs(n, m, ptr, col, cp, row, int(r))
        assert np.array_equal(outs["graph"][i], ref), i
    groups = -(-roots.size // lanes)
    for form, info in infos.items():
        assert info["traversals"] == groups and info["lanes"] == lanes
        assert info["fallbacks"] == 0
    # the same levels, decided on the device or on the host
    assert infos["graph"]["levels"] == infos["host"]["levels"]
    assert infos["graph"]["push_levels"] == infos["host"]["push_levels"]
    if not undirected:
        assert infos["graph"]["push_levels"] == 0  # directed: pull levels only
    # host waits: two per level host-driven, one per batch of levels otherwise
    assert infos["host"]["host_syncs"] == 2 * infos["host"]["levels"]
    assert infos["graph"]["host_syncs"] <= groups * (1 + infos["graph"]["levels"] // 4)
    assert infos["graph"]["host_syncs"] < infos["graph"]["levels"]


def test_deep_path_falls_back_to_single_source():
    """A path of 70,000 vertices is deeper than the 16-bit level store: the
    traversal stops there and each source reruns on the single-source path,
    in the batch call and in the serving engine's lane groups."""
    n = 70_000
    v = np.arange(n - 1, dtype=np.uint32)
    roots = np.array([0, n - 1, 12_345], np.uint32)
    with eg.Graph(n, eg.MODE_ADJ_SORT, undirected=True, symmetric=True) as g:
        g.load_edges(v, v + 1)
        print("deep path: batched", flush=True)
        out, _ = g.bfs_batch(roots)
        info = g.bfs_batch_info()
        print("deep path: serving engine", flush=True)
        with g.serve(idle_plain=-1) as S:
            t = [S.submit_bfs(int(r)) for r in roots]
            res = [S.wait(x) for x in t]
    assert info["fallbacks"] == 1
    ids = np.arange(n, dtype=np.int64)
    for i, r in enumerate(roots):
        ref = (np.abs(ids - int(r)) + 1).astype(np.uint32)  # level + 1 on a path
        assert np.array_equal(out[i], ref), i
        r_info, r_out = res[i]
        assert np.array_equal(r_out, ref), i
        assert r_info["reached"] == n
