"""The brick(...) / tile(...) operator hooks and multiarray.h on the host.

* The reference's own external example (examples/external/main.cpp of
  bricklib), unchanged, compiles against include/ as C++14 (its CMake
  standard) and runs to completion: its two compareBrick checks -- the
  accessor brick loop and the brick(...) loop against the array loop and
  the tile(...) loop -- pass (it throws "result mismatch!" otherwise).
  Skipped where the reference tree is absent.
* tests/cpp/hook_check.cpp sweeps host bricks through brick("7pt.py", ...)
  and brick("cond.py", ...): the bricks equal the oracle's fixed-order
  sweep bit for bit, and the tile(...) loop equals them within
  compareBrick's tolerance.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_EXAMPLE = "/root/reference/examples/external/main.cpp"


def gxx(src, out, std="c++17", extra=()):
    subprocess.run(["g++", "-O2", f"-std={std}", "-fopenmp", "-Wall", "-Werror", "-I",
                    os.path.join(ROOT, "include"), *extra, src, "-o", out], check=True)


@pytest.mark.skipif(not os.path.exists(REF_EXAMPLE), reason="reference tree not present")
def test_reference_external_example_compiles_and_runs(tmp_path):
    exe = str(tmp_path / "external")
    gxx(REF_EXAMPLE, exe, std="c++14")
    env = dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1)))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "d3pt7" in r.stdout and "Trans:" in r.stdout


@pytest.fixture(scope="module")
def hook_exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hook") / "hook_check")
    gxx(os.path.join(ROOT, "tests", "cpp", "hook_check.cpp"), out, std="c++14")
    return out


@pytest.mark.parametrize("stencil,kind", [("7pt.py", O.STAR7), ("cond.py", O.COND7)])
@pytest.mark.parametrize("nb", [2, 3])
def test_host_hook_matches_oracle(hook_exe, tmp_path, stencil, kind, nb):
    path = str(tmp_path / "bricks.bin")
    r = subprocess.run([hook_exe, str(nb), stencil, path], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "tile == brick" in r.stdout
    g = O.Grid((nb, nb, nb))
    raw = np.fromfile(path, np.float64)
    dat = raw[:g.nbricks * 1024].reshape(g.nbricks, 1024)
    coeff = raw[g.nbricks * 1024:]
    want = np.zeros_like(dat)
    want[:, :512] = dat[:, :512]
    g.sweep(kind, coeff, want, want, in_off=0, out_off=512, cond_mode=0)
    np.testing.assert_array_equal(dat[g.interior, 512:], want[g.interior, 512:])


def test_unknown_stencil_file_throws(tmp_path):
    src = tmp_path / "bad.cpp"
    src.write_text('#include "brick.h"\n#include "bricksetup.h"\nbElem coeff[7];\n'
                   "int main() {\n  unsigned *g;\n  auto info = init_grid<3>(g, {3, 3, 3});\n"
                   "  auto st = BrickStorage::allocate(info.nbricks, 1024);\n"
                   "  Brick<Dim<8,8,8>, Dim<8,8>> bIn(&info, st, 0), bOut(&info, st, 512);\n"
                   "  try { brick(\"mine.py\", \"X\", (8,8,8), (8,8), 13); }\n"
                   "  catch (const std::invalid_argument &) { return 0; }\n  return 1;\n}\n")
    exe = str(tmp_path / "bad")
    gxx(str(src), exe)
    assert subprocess.run([exe]).returncode == 0
