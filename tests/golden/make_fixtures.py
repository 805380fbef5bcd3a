"""Writes the committed R-MAT 10 fixtures of the test suite.

The reference ships no test data (no golden files, no datasets), so these
fixtures pin this repository's own generator and CPU oracle against later
drift: a change to either shows up as a mismatch in tests/test_golden.py.

  rmat10.bin          the input: 16,384 records of R-MAT 10 (edge factor 16,
                      a, b, c = 0.57, 0.19, 0.19, seed 1) in the reference's
                      binary [src][dst] uint32 little-endian record format
                      (init_all.h:8-14), self-loops and duplicates kept
  rmat10_expected.npz what the oracle computes on it:
      csr_ptr, csr_col      CSR of load mode 2 (init_adj_sort, init_all.c:629-697)
      csc_ptr, csc_row      CSC of the same graph (init_all.c:714-767)
      u_csr_ptr, u_csr_col  CSR with -u (reversed copies at i + m_in,
                            init_all.c:102-106)
      grid_ptr, grid_src, grid_dst  2-D grid, P = 4, mode 0 (init_grid_sort)
      bfs_roots, bfs_levels  BFS levels (root 1, unreached 0) on the -u graph
                            from four roots (bfs_simple.c levels)
      wcc_labels            min-label components of the -u graph (wcc.c)
      pagerank              FP64 pull PageRank of the directed graph, d = 0.85,
                            20 iterations, r0 = 1/n (pagerank_simple.c:72-84)
  MANIFEST.json       parameters and the SHA-256 of both files

Run from the repository root (needs oracle/liboracle.so):
    python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import eg_oracle_py as O  # noqa: E402

SCALE = 10
EDGE_FACTOR = 16
SEED = 1
GRID_P = 4
ROOTS = (0, 1, 2, 3)  # the k-th vertex with edges, k in ROOTS


def expected(src, dst):
    """Every fixture array, from the oracle, for records (src, dst)."""
    n = 1 << SCALE
    ptr, col, _ = O.build_csr(n, src, dst)
    cp, row, _ = O.build_csc(n, ptr, col)
    us, ud, _ = O.symmetrize(src, dst)
    uptr, ucol, _ = O.build_csr(n, us, ud)
    ucp, urow, _ = O.build_csc(n, uptr, ucol)
    gp, gs, gd, _ = O.build_grid(n, GRID_P, src, dst)
    with_edges = np.flatnonzero(np.diff(uptr) > 0)
    roots = with_edges[list(ROOTS)].astype(np.uint32)
    levels = np.stack([O.bfs(n, us.size, uptr, ucol, ucp, urow, int(r))[0] for r in roots])
    labels, _ = O.wcc_push(n, uptr, ucol)
    rank, _, _ = O.pagerank_pull(n, cp, row, np.diff(ptr).astype(np.uint32), 0.85, 20)
    return {
        "csr_ptr": ptr, "csr_col": col, "csc_ptr": cp, "csc_row": row,
        "u_csr_ptr": uptr, "u_csr_col": ucol,
        "grid_ptr": gp, "grid_src": gs, "grid_dst": gd,
        "bfs_roots": roots, "bfs_levels": levels, "wcc_labels": labels, "pagerank": rank,
    }


def records():
    return O.gen_rmat(SCALE, EDGE_FACTOR << SCALE, seed=SEED)


def sha256(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def main():
    src, dst = records()
    rec = np.empty((src.size, 2), "<u4")
    rec[:, 0], rec[:, 1] = src, dst
    bin_path = os.path.join(HERE, "rmat10.bin")
    npz_path = os.path.join(HERE, "rmat10_expected.npz")
    rec.tofile(bin_path)
    np.savez_compressed(npz_path, **expected(src, dst))
    manifest = {
        "graph": {"kind": "rmat", "scale": SCALE, "edge_factor": EDGE_FACTOR, "seed": SEED,
                  "abc": list(O.RMAT_ABC), "records": int(src.size)},
        "grid_P": GRID_P, "bfs_roots": "k-th vertex with edges in the -u graph, k in "
                                       + ",".join(map(str, ROOTS)),
        "pagerank": {"damping": 0.85, "iterations": 20, "init": "1/n", "dangling": "drop"},
        "sha256": {"rmat10.bin": sha256(bin_path), "rmat10_expected.npz": sha256(npz_path)},
    }
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1)
        f.write("\n")
    print(json.dumps(manifest["sha256"], indent=1))


if __name__ == "__main__":
    main()
