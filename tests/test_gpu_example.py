"""The C++ drop-in example (examples/drop_in_7pt.cpp) on the GPU: bricklib
user code -- init_grid, copyToBrick, movBrickInfo/movBrickStorage, the
sweep, compareBrick -- against its own host array loop, 1 sweep and 8
ping-pong sweeps with the on-device periodic ghost refresh."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "examples", "bin", "drop_in_7pt")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,sweeps", [(32, 1), (64, 8), (48, 5)])
def test_drop_in_example(gpu, n, sweeps):
    assert os.path.exists(EXE), "examples/bin/drop_in_7pt not built (run build())"
    r = subprocess.run([EXE, str(n), str(sweeps)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "result match" in r.stdout


DECOMP_EXE = os.path.join(ROOT, "examples", "bin", "decomp_7pt")


@pytest.mark.parametrize("n,sweeps,skin", [(32, 4, "good"), (24, 3, "normal"), (16, 5, "bad")])
def test_decomp_example(gpu, n, sweeps, skin):
    """BrickDecomp (include/brick-decomp.h) on device storage: exchange() +
    brick_sweep per step, and brick_sweeps over device_pairs(), against the
    host array loop; the two device paths bit-identical."""
    assert os.path.exists(DECOMP_EXE), "examples/bin/decomp_7pt not built (run build())"
    print(f"[case] decomp_7pt n={n} sweeps={sweeps} skin={skin}", flush=True)
    r = subprocess.run([DECOMP_EXE, str(n), str(sweeps), skin], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "bit-identical" in r.stdout and "result match" in r.stdout


HOOK_EXE = os.path.join(ROOT, "examples", "bin", "brick_hook_7pt")


@pytest.mark.parametrize("n,sweeps", [(32, 1), (64, 4), (40, 3)])
def test_brick_hook_example(gpu, n, sweeps):
    """examples/brick_hook_7pt.cpp keeps the reference's per-brick
    brick("7pt.py", ...) call site: on host bricks it computes at once, on
    device-resident bricks it queues them and runs one GPU sweep at the next
    flush; both against the array loop (compareBrick) and against each other
    bit for bit, over ping-pong sweeps too; tile(...) against the array
    loop."""
    assert os.path.exists(HOOK_EXE), "examples/bin/brick_hook_7pt not built (run build())"
    print(f"[case] brick_hook_7pt n={n} sweeps={sweeps}", flush=True)
    r = subprocess.run([HOOK_EXE, str(n), str(sweeps)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "result match" in r.stdout
