"""The C++ BrickDecomp drop-in (include/brick-decomp.h) on the host: BitSet
algebra, the numbering it builds for each of the reference's surface orders
(skin3d_good / normal / bad) against bk_decomp_layout(_skin) as the Python
side reads it -- grid, sep_pos, skin_size, every ghost/skin run with its
neighbour -- and the host exchange: each ghost element equals its periodic
image afterwards, for two interleaved fields."""
import os
import subprocess

import numpy as np
import pytest

from bricklib_amd.decomp import SKIN3D_BAD, SKIN3D_GOOD, SKIN3D_NORMAL, DecompLayout

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "bricklib_amd")
SKINS = {"good": SKIN3D_GOOD, "normal": SKIN3D_NORMAL, "bad": SKIN3D_BAD}


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    assert os.path.exists(os.path.join(LIBDIR, "libbricks.so")), "run build() first"
    out = str(tmp_path_factory.mktemp("cpp") / "decomp_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I",
                    os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "decomp_check.cpp"), "-L", LIBDIR,
                    "-lbricks", f"-Wl,-rpath,{LIBDIR}", "-o", out], check=True)
    return out


def run(exe, *args):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_bitset(exe):
    assert "bitset ok" in run(exe, "bitset")


@pytest.mark.parametrize("skin", sorted(SKINS))
@pytest.mark.parametrize("dims,numfield,B", [((3, 4, 5), 1, 8), ((2, 2, 2), 2, 8),
                                             ((4, 3, 2), 1, 4), ((3, 3, 3), 2, 4)])
def test_layout_matches_c_abi(exe, skin, dims, numfield, B):
    out = run(exe, "layout", skin, *dims, numfield, B)
    brick_bytes = B ** 3 * 8 * numfield
    L = DecompLayout(dims, 1, brick_bytes=brick_bytes, page_bytes=os.sysconf("SC_PAGESIZE"),
                     skin=SKINS[skin])
    runs = []
    for line in out.splitlines():
        f = line.split()
        if f[0] == "sep":
            assert tuple(map(int, f[1:])) == L.sep
        elif f[0] == "skin_size":
            assert list(map(int, f[1:])) == L.skin_size
        elif f[0] == "run":
            runs.append(tuple(map(int, f[1:])))
        elif f[0] == "grid":
            np.testing.assert_array_equal(np.array(f[1:], np.uint32), L.grid.reshape(-1))
    want = [tuple(r["neighbor"]) + (r["ghost_pos"], r["skin_pos"], r["len"], r["skin_st"],
                                     r["skin_ed"], r["first_pad"], r["last_pad"])
            for r in L.regions]
    assert runs == want


@pytest.mark.parametrize("skin,dims", [("good", (3, 4, 5)), ("normal", (4, 2, 3)),
                                       ("bad", (2, 2, 3)), ("good", (2, 2, 2))])
def test_host_exchange_fills_periodic_images(exe, skin, dims):
    out = run(exe, "exchange", skin, *dims)
    n = np.prod([d + 2 for d in dims]) - np.prod(dims)
    assert f"checked {n} ghost bricks" in out and "exchange ok" in out
