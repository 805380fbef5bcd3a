"""Layout builders on the device vs the oracle: bit-exact CSR / CSC / grid /
edge arrays for every load mode, the generator, file ingest and error paths."""
import os

import numpy as np
import pytest

import eg_oracle_py as O
import everythinggraph_amd as eg

pytestmark = pytest.mark.gpu


def graphs():
    out = []
    out.append(("path", 6, np.arange(5, dtype=np.uint32), np.arange(1, 6, dtype=np.uint32)))
    out.append(("self_dup", 4, np.array([0, 0, 1, 1, 3, 3], np.uint32),
                np.array([0, 1, 2, 2, 3, 0], np.uint32)))
    out.append(("isolated_tail", 5000, np.array([3, 7, 3, 9], np.uint32),
                np.array([7, 3, 9, 3], np.uint32)))
    rng = np.random.default_rng(5)
    s = rng.integers(0, 3000, 60000).astype(np.uint32)
    s[:25000] = 11  # hub well past one tile and one long-row segment
    d = rng.integers(0, 3000, 60000).astype(np.uint32)
    out.append(("hub", 3000, s, d))
    s, d = O.gen_rmat(12, 16 << 12, seed=1)
    out.append(("rmat12", 1 << 12, s, d))
    return out


GRAPHS = graphs()
IDS = [g[0] for g in GRAPHS]


def stored(s, d, w, undirected):
    if undirected:
        return O.symmetrize(s, d, w)
    return s, d, w


@pytest.mark.parametrize("name,n,s,d", GRAPHS, ids=IDS)
@pytest.mark.parametrize("undirected", [False, True])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("mode", [2, 3, 4])
def test_adjacency_bit_exact(name, n, s, d, undirected, weighted, mode):
    w = O.gen_weights(s.size, seed=9) if weighted else None
    with eg.Graph(n, mode, undirected=undirected, weighted=weighted) as g:
        g.load_edges(s, d, w)
        ss, sd, sw = stored(s, d, w, undirected)
        ptr, col, ow = O.build_csr(n, ss, sd, sw)
        cp, row, iw = O.build_csc(n, ptr, col, ow)
        gptr, gcol, gw = g.csr()
        assert np.array_equal(gptr, ptr) and np.array_equal(gcol, col)
        if weighted:
            assert np.array_equal(gw, ow)
        hp, hr, hw = g.csc()
        assert np.array_equal(hp, cp) and np.array_equal(hr, row)
        if weighted:
            assert np.array_equal(hw, iw)
        od, idg = g.degrees()
        assert np.array_equal(od, np.diff(ptr).astype(np.uint32))
        assert np.array_equal(idg, np.diff(cp).astype(np.uint32))
        inf = g.info
        assert inf.m == ss.size and inf.m_in == s.size and inf.has_csr and inf.has_csc


def test_symmetric_flag_skips_csc():
    s, d = O.gen_rmat(8, 2048, seed=2)
    with eg.Graph(256, 2, undirected=True, symmetric=True) as g:
        g.load_edges(s, d)
        assert not g.info.has_csc
        with pytest.raises(eg.EgError) as e:
            g.csc()
        assert e.value.code == eg.EG_ESTATE


@pytest.mark.parametrize("name,n,s,d", GRAPHS, ids=IDS)
@pytest.mark.parametrize("mode", [0, 1, 6])
@pytest.mark.parametrize("Pp", [1, 5, 16])
def test_grid_bit_exact(name, n, s, d, mode, Pp):
    w = O.gen_weights(s.size, seed=4)
    with eg.Graph(n, mode, undirected=True, weighted=True, P=Pp) as g:
        g.load_edges(s, d, w)
        ss, sd, sw = O.symmetrize(s, d, w)
        cp, gs, gd, gw = O.build_grid(n, Pp, ss, sd, sw, presort=(mode != 1))
        hcp, hs, hd, hw = g.grid()
        assert np.array_equal(hcp, cp)
        assert np.array_equal(hs, gs) and np.array_equal(hd, gd) and np.array_equal(hw, gw)


def test_grid_default_P():
    s, d = O.gen_rmat(16, 1 << 16, seed=3)
    with eg.Graph(1 << 16, 0) as g:
        g.load_edges(s, d)
        assert g.info.P == O.default_P(1 << 16) == 25
        cp, gs, gd, _ = O.build_grid(1 << 16, 25, s, d)
        hcp, hs, hd, _ = g.grid()
        assert np.array_equal(hcp, cp) and np.array_equal(hs, gs) and np.array_equal(hd, gd)


@pytest.mark.parametrize("name,n,s,d", GRAPHS, ids=IDS)
@pytest.mark.parametrize("mode", [7, 8, 9])
def test_edgelist_bit_exact(name, n, s, d, mode):
    w = O.gen_weights(s.size, seed=6)
    with eg.Graph(n, mode, weighted=True) as g:
        g.load_edges(s, d, w)
        os_, od, ow = O.build_edgelist(n, s, d, w, mode)
        hs, hd, hw = g.edgelist()
        assert np.array_equal(hs, os_) and np.array_equal(hd, od) and np.array_equal(hw, ow)


def test_generator_matches_oracle():
    p = eg.rmat_params(12, 16, seed=5)
    s, d, w = eg.generate_edges(p, 1 << 12, weighted=True)
    os_, od = O.gen_rmat(12, 16 << 12, seed=5)
    assert np.array_equal(s, os_) and np.array_equal(d, od)
    assert np.array_equal(w, O.gen_weights(16 << 12, seed=1))
    p = eg.er_params(100000, seed=7, weight_seed=3)
    s, d, w = eg.generate_edges(p, 77777, weighted=True)
    os_, od = O.gen_er(77777, 100000, seed=7)
    assert np.array_equal(s, os_) and np.array_equal(d, od)
    assert np.array_equal(w, O.gen_weights(100000, seed=3))


def test_generate_in_place_equals_loaded():
    p = eg.rmat_params(11, 8, seed=2)
    n = 1 << 11
    with eg.Graph(n, 2, undirected=True) as a, eg.Graph(n, 2, undirected=True) as b:
        a.generate(p)
        s, d, _ = eg.generate_edges(p, n)
        b.load_edges(s, d)
        for x, y in zip(a.csr()[:2], b.csr()[:2]):
            assert np.array_equal(x, y)


def test_file_ingest(tmp_path):
    s, d = O.gen_rmat(10, 5000, seed=8)
    w = O.gen_weights(5000, seed=8)
    rec = np.empty((5000, 2), np.uint32)
    rec[:, 0], rec[:, 1] = s, d
    f = tmp_path / "g.bin"
    # a trailing partial record is dropped (st_size / item_size, init_all.c:821)
    f.write_bytes(rec.tobytes() + b"\x01\x02\x03")
    with eg.Graph(1024, 2) as g:
        g.load_file(str(f))
        assert g.info.m_in == 5000
        ptr, col, _ = O.build_csr(1024, s, d)
        assert np.array_equal(g.csr()[1], col)
    recw = np.empty((5000, 3), np.uint32)
    recw[:, 0], recw[:, 1], recw[:, 2] = s, d, w.view(np.uint32)
    fw = tmp_path / "gw.bin"
    fw.write_bytes(recw.tobytes())
    with eg.Graph(1024, 8, weighted=True) as g:
        g.load_file(str(fw))
        os_, od, ow = O.build_edgelist(1024, s, d, w, 8)
        hs, hd, hw = g.edgelist()
        assert np.array_equal(hs, os_) and np.array_equal(hd, od) and np.array_equal(hw, ow)
    with eg.Graph(1024, 2) as g:
        g.load_records(rec.tobytes())
        assert np.array_equal(g.csr()[1], col)


@pytest.mark.parametrize("kib", [1, 3, 7])
def test_chunked_ingest(tmp_path, kib):
    """The pinned two-slot ingest pipeline with chunks far smaller than the
    file (many chunks, a ragged last chunk, 12-byte records that do not
    divide the chunk) builds the same layouts as one-shot loads."""
    n = 1 << 11
    s, d = O.gen_rmat(11, 20011, seed=9)
    w = O.gen_weights(s.size, seed=9)
    recw = np.empty((s.size, 3), np.uint32)
    recw[:, 0], recw[:, 1], recw[:, 2] = s, d, w.view(np.uint32)
    fw = tmp_path / "gw.bin"
    fw.write_bytes(recw.tobytes() + b"\x07" * 11)  # partial trailing record
    rec = np.ascontiguousarray(recw[:, :2])
    f = tmp_path / "g.bin"
    f.write_bytes(rec.tobytes())
    ptr, col, ow = O.build_csr(n, *O.symmetrize(s, d, w))
    with eg.Graph(n, 2, undirected=True, weighted=True) as g:
        g.tune_set("ingest_kib", kib)
        g.load_file(str(fw))
        assert g.info.m_in == s.size
        hp, hc, hw = g.csr()
        assert np.array_equal(hp, ptr) and np.array_equal(hc, col) and np.array_equal(hw, ow)
    ptr2, col2, _ = O.build_csr(n, s, d)
    with eg.Graph(n, 2) as g:
        g.tune_set("ingest_kib", kib)
        g.load_file(str(f))
        assert np.array_equal(g.csr()[1], col2)
        g.load_records(rec.tobytes())  # reload from host memory through the same pipeline
        assert np.array_equal(g.csr()[1], col2)


def test_error_paths(tmp_path):
    with pytest.raises(eg.EgError) as e:
        eg.Graph(10, 5)
    assert e.value.code == eg.EG_ENOTSUP
    with pytest.raises(eg.EgError) as e:
        eg.Graph(10, 11)
    assert e.value.code == eg.EG_EINVAL
    with eg.Graph(10, 2) as g:
        with pytest.raises(eg.EgError) as e:
            g.load_edges(np.array([1, 10], np.uint32), np.array([2, 3], np.uint32))
        assert e.value.code == eg.EG_ERANGE
        with pytest.raises(eg.EgError) as e:
            g.load_file(str(tmp_path / "missing.bin"))
        assert e.value.code == eg.EG_EIO
        with pytest.raises(eg.EgError) as e:
            g.pagerank()
        assert e.value.code == eg.EG_ESTATE
    with eg.Graph(10, 2) as g:
        g.load_edges(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
        ptr, col, _ = g.csr()
        assert np.all(ptr == 0) and col.size == 0
        with pytest.raises(eg.EgError) as e:
            g.bfs(10)
        assert e.value.code == eg.EG_ERANGE


def test_determinism_repeat_build():
    s, d = O.gen_rmat(13, 16 << 13, seed=3)
    outs = []
    for _ in range(2):
        with eg.Graph(1 << 13, 2, undirected=True) as g:
            g.load_edges(s, d)
            outs.append(g.csc())
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("undirected", [False, True])
def test_multi_tile_ranges_bit_exact(undirected):
    """Past 2048 tiles of 8192 records a sort range holds several tiles (run
    offsets cross tiles and ranges), and the last vertices have no entries."""
    n = (1 << 20) + 777
    s, d = O.gen_rmat(20, 20 << 20, seed=17)
    with eg.Graph(n, 2, undirected=undirected) as g:
        g.load_edges(s, d)
        ss, sd, _ = stored(s, d, None, undirected)
        assert ss.size > 2048 * 8192
        ptr, col, _ = O.build_csr(n, ss, sd)
        cp, row, _ = O.build_csc(n, ptr, col)
        gptr, gcol, _ = g.csr()
        assert np.array_equal(gptr, ptr) and np.array_equal(gcol, col)
        hp, hr, _ = g.csc()
        assert np.array_equal(hp, cp) and np.array_equal(hr, row)
        od, idg = g.degrees()
        assert np.array_equal(od, np.diff(ptr)) and np.array_equal(idg, np.diff(cp))


def test_empty_and_single_record():
    for s in (np.zeros(0, np.uint32), np.array([41], np.uint32)):
        with eg.Graph(64, 2) as g:
            g.load_edges(s, s)
            ptr, col, _ = g.csr()
            assert np.array_equal(ptr, np.concatenate([[0], np.cumsum(np.bincount(s, minlength=64))]))
            assert np.array_equal(col, s)
