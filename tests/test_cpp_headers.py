"""The C++ drop-in headers (include/brick.h, bricksetup.h, brickcompare.h)
on the host: grid numbering/adjacency, accessor offsets for several vector
folds and through neighbours, and the copyToBrick/copyFromBrick round trip
with compareBrick -- against the oracle and the reference's own headers
(oracle/_ref)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cpp") / "header_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fopenmp", "-Wall", "-Werror", "-I",
                    os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp",
                                                                "header_check.cpp"),
                    "-o", out], check=True)
    return out


def run(exe, *args):
    return subprocess.run([exe, *map(str, args)], check=True, capture_output=True,
                          text=True).stdout.split("\n")


@pytest.mark.parametrize("dims,periodic", [((3, 4, 5), 0), ((1, 1, 1), 0), ((4, 3, 2), 1),
                                           ((6, 6, 6), 0)])
def test_grid_matches_oracle(exe, dims, periodic):
    vals = np.array([int(x) for x in run(exe, "grid", *dims, periodic) if x], np.uint32)
    n = int(np.prod(dims))
    g, a = O.init_grid(dims, periodic=bool(periodic))
    np.testing.assert_array_equal(vals[:n], g)
    np.testing.assert_array_equal(vals[n:].reshape(n, 27), a)


@pytest.mark.parametrize("fold", range(7))
def test_fold_offsets_match_reference(exe, fold):
    got = np.array([int(x) for x in run(exe, "fold", fold) if x])
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    grid, adj = O.init_grid((3, 3, 3))
    exp = np.empty(512, np.int64)
    O.ref().ref_fold_offsets(fold, O.P(adj), adj.shape[0], 13, C.c_size_t(512), O.P(exp))
    np.testing.assert_array_equal(got, exp)


def test_neighbor_offsets_match_reference(exe):
    lines = [x for x in run(exe, "neighbor") if x]
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    R = O.ref()
    R.ref_neighbor_offset.restype = C.c_long
    R.ref_neighbor_offset.argtypes = [C.c_void_p, C.c_uint, C.c_uint, C.c_size_t, C.c_int,
                                      C.c_int, C.c_int]
    grid, adj = O.init_grid((3, 3, 3))
    for ln in lines[:-1]:
        k, j, i, off = (int(v) for v in ln.split())
        assert R.ref_neighbor_offset(O.P(adj), adj.shape[0], 13, 512, k, j, i) == off
    # neighbor<2,1,0>(13): digits k=2, j=1, i=0 -> slot 21
    assert lines[-1] == f"nb {int(adj[13][21]) * 512}"


def test_copy_round_trip_and_compare(exe):
    out = run(exe, "copy")[0]
    assert out == "touched 3072 diff 0 same 1 caught 1"
