"""The drop-in binaries end to end: each reads an edge file with the
reference's CLI (random.c:50-116), runs on the device and prints the lines of
the reference binary; the printed results are checked against the oracle.

With EG_DUMP set, each binary also writes the whole result array it leaves
behind (dist, components, rank, y) in the reference encoding, and that array
is compared with the oracle: bit for bit for levels, labels and distances,
within the FP32 tolerances of the algorithm tests for ranks and products."""
import os
import re
import subprocess

import numpy as np
import pytest

import eg_oracle_py as O
import everythinggraph_amd as eg

pytestmark = pytest.mark.gpu

N = 1 << 10


@pytest.fixture(scope="module")
def graph_file(tmp_path_factory):
    s, d = O.gen_rmat(10, 8 << 10, seed=77)
    rec = np.empty((s.size, 2), np.uint32)
    rec[:, 0], rec[:, 1] = s, d
    path = tmp_path_factory.mktemp("drv") / "rmat10.bin"
    path.write_bytes(rec.tobytes())
    return str(path), s, d


def run_dump(tmp_path, dtype, drv, *args, env=None):
    """(stdout, the result array the binary wrote through EG_DUMP)."""
    path = tmp_path / f"{drv}.out"
    exe = os.path.join(eg.BIN_DIR, drv)
    print("driver:", drv, *args, env or "", flush=True)
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=120,
                       env={**os.environ, "EG_DUMP": str(path), **(env or {})})
    assert r.returncode == 0, (drv, args, r.stdout[-2000:], r.stderr[-2000:])
    out = np.fromfile(path, dtype)
    assert out.size == N
    return r.stdout, out


def ranks_printed(out):
    return {int(i): float(v) for i, v in re.findall(r"Rank\[(\d+)\] = ([0-9.eE+-]+)", out)}


@pytest.mark.parametrize("drv,mode", [("pagerank_simple", 2), ("pagerank_simple", 8),
                                      ("prgrid_cilk", 0)])
def test_pagerank_drivers(graph_file, tmp_path, drv, mode):
    path, s, d = graph_file
    out, rank = run_dump(tmp_path, np.float32, drv, "-f", path, "-n", N, "-m", mode, "-i", 20,
                         "-I", 0)
    ptr, col, _ = O.build_csr(N, s, d)
    cp, row, _ = O.build_csc(N, ptr, col)
    ref, _, _ = O.pagerank_pull(N, cp, row, np.diff(ptr).astype(np.uint32), 0.85, 20)
    got = ranks_printed(out)
    assert sorted(got) == list(range(10))
    for i, v in got.items():
        assert abs(v - ref[i]) <= 6e-6, (i, v, ref[i])  # printed with 5 decimals
    err = np.abs(rank.astype(np.float64) - ref)
    assert err.sum() <= 1e-6 and err.max() <= 1e-6 * ref.max()


def test_spmv_driver(graph_file, tmp_path):
    path, s, d = graph_file
    out, yall = run_dump(tmp_path, np.float32, "spmv", "-f", path, "-n", N, "-m", 2)
    y = {int(i): float(v) for i, v in re.findall(r"y\[(\d+)\] = ([0-9.eE+-]+)", out)}
    ref = np.zeros(N)
    np.add.at(ref, d, s.astype(np.float64))  # y[v] = sum over u->v of x[u], x[i] = i
    for i, v in y.items():
        assert abs(v - ref[i]) <= 1e-5 * max(1.0, ref[i]) + 6e-6, (i, v, ref[i])
    total = float(re.search(r"sum\(y\) = ([0-9.eE+-]+)", out).group(1))
    assert abs(total - ref.sum()) <= 1e-5 * ref.sum()
    # x >= 0 here, so |A| |x| = A x
    assert np.all(np.abs(yall.astype(np.float64) - ref) <= 1e-5 * ref + 1e-30)


@pytest.mark.parametrize("drv,mode", [("bfs_simple", 2), ("bfs_simple", 4),
                                      ("bfsgrid_cilk", 0)])
def test_bfs_drivers(graph_file, tmp_path, drv, mode):
    path, s, d = graph_file
    root = int(s[0])
    out, dist = run_dump(tmp_path, np.uint32, drv, "-f", path, "-n", N, "-m", mode, "-u", "-p",
                         root)
    su, du, _ = O.symmetrize(s, d)
    ptr, col, _ = O.build_csr(N, su, du)
    cp, row, _ = O.build_csc(N, ptr, col)
    ref, _, _ = O.bfs(N, su.size, ptr, col, cp, row, root)
    found = int(re.search(r"Total nodes discovered:: (\d+)", out).group(1))
    assert found == np.count_nonzero(ref)
    assert np.array_equal(dist, ref)  # level + 1, 0 = unreached


@pytest.mark.parametrize("mode", [2, 4])
def test_bfs_driver_wide_offsets(graph_file, tmp_path, mode):
    """bfs_simple on the layout with 64-bit row offsets (what a load past 2^32
    stored entries builds), forced on a small graph through EG_TUNE: parts of
    2^12 records, so the rows are placed across several sorted parts."""
    path, s, d = graph_file
    root = int(s[0])
    out, dist = run_dump(tmp_path, np.uint32, "bfs_simple", "-f", path, "-n", N, "-m", mode,
                         "-u", "-p", root, env={"EG_TUNE": "wide_part=12"})
    su, du, _ = O.symmetrize(s, d)
    ptr, col, _ = O.build_csr(N, su, du)
    cp, row, _ = O.build_csc(N, ptr, col)
    ref, _, _ = O.bfs(N, su.size, ptr, col, cp, row, root)
    assert np.array_equal(dist, ref)


@pytest.mark.parametrize("tune", ["wide_part", "no_such_knob=1", "pr_binned=x"])
def test_driver_rejects_bad_tune_env(graph_file, tune):
    path, _, _ = graph_file
    exe = os.path.join(eg.BIN_DIR, "bfs_simple")
    r = subprocess.run([exe, "-f", path, "-n", str(N), "-u"], capture_output=True, text=True,
                       timeout=120, env={**os.environ, "EG_TUNE": tune})
    assert r.returncode != 0 and "EG_TUNE" in r.stderr


@pytest.mark.parametrize("mode", [2, 8, 0])
def test_wcc_driver(graph_file, tmp_path, mode):
    path, s, d = graph_file
    out, comp = run_dump(tmp_path, np.int64, "wcc", "-f", path, "-n", N, "-m", mode, "-u")
    su, du, _ = O.symmetrize(s, d)
    ptr, col, _ = O.build_csr(N, su, du)
    lab, _ = O.wcc_push(N, ptr, col)
    sizes = np.bincount(lab, minlength=N)
    tc, mc = map(int, re.search(r"Total comp (\d+) , max comp (\d+)", out).groups())
    assert tc == np.count_nonzero(sizes) and mc == sizes.max()
    assert np.array_equal(comp, lab)  # long components[] of wcc.c:12


def test_sssp_driver(graph_file, tmp_path):
    path, s, d = graph_file
    root = int(s[0])
    out, dist = run_dump(tmp_path, np.uint32, "sssp_pushpull", "-f", path, "-n", N, "-m", 2,
                         "-p", root)
    ptr, col, _ = O.build_csr(N, s, d)
    ref, _ = O.sssp(N, ptr, col, root, 0)
    fin = ref[ref != O.SSSP_UNREACHED]
    mx, d1 = map(int, re.search(r"Max dist = (\d+) and dist\[1\] is (-?\d+)", out).groups())
    assert mx == fin.max()
    # dist[1] is printed as a signed int: the unreached marker UINT32_MAX - 1 reads -2
    assert d1 == (int(ref[1]) if ref[1] != O.SSSP_UNREACHED else -2)
    found = int(re.search(r"Total nodes discovered:: (\d+)", out).group(1))
    assert found == fin.size
    assert np.array_equal(dist, ref)
