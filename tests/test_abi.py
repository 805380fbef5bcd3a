"""Host-side checks that need no GPU: the C-ABI library loads and exports
every function include/eg.h declares; the drop-in drivers parse the
reference's command line; the library reports a missing device loudly."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import everythinggraph_amd as eg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "eg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(eg_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_python_symbol_list():
    assert declared_functions() == sorted(eg.ABI_SYMBOLS)


def test_library_exports_every_declared_symbol():
    L = eg.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", eg.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert exported == set(declared_functions())


def test_version_string():
    assert b"gfx950" in eg.lib().eg_version()


DRIVERS = ["pagerank_simple", "prgrid_cilk", "spmv", "bfs_simple", "bfsgrid_cilk", "wcc",
           "sssp_pushpull"]


@pytest.mark.parametrize("drv", DRIVERS)
def test_driver_usage(drv):
    exe = os.path.join(eg.BIN_DIR, drv)
    r = subprocess.run([exe, "-h"], capture_output=True, text=True)
    assert r.returncode != 0 and "Usage: ./random -f <graph file> -n <nb_nodes>" in r.stdout
    r = subprocess.run([exe, "-n", "10"], capture_output=True, text=True)  # no -f
    assert r.returncode != 0 and "Usage" in r.stdout
    r = subprocess.run([exe, "-f", "x", "-n", "10", "-q"], capture_output=True, text=True)
    assert r.returncode != 0 and "Unknown option" in r.stderr


def _has_gpu():
    d = eg.DeviceInfo()
    return eg.lib().eg_device_query(ctypes.byref(d)) == 0


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_no_device_is_an_error():
    with pytest.raises(eg.EgError) as e:
        eg.Graph(16, 2)
    assert e.value.code == eg.EG_EDEVICE
    with pytest.raises(eg.EgError):
        eg.generate_edges(eg.rmat_params(4, 2), 16)


def test_sample_roots_properties():
    deg = np.zeros(1000, np.int64)
    deg[::3] = 1
    r = eg.sample_roots(deg, 50, seed=2)
    assert len(set(r.tolist())) == 50 and np.all(deg[r] > 0)
    assert np.array_equal(r, eg.sample_roots(deg, 50, seed=2))
    assert not np.array_equal(r, eg.sample_roots(deg, 50, seed=3))
