"""BF16 row-form kernels (one brick per wave, and 2x2 blocks of bricks)
against the oracle on seeded random cases: grid extents 1..5 bricks per axis
(one-brick axes make a brick its own neighbour; odd extents leave partial
blocks and levels), every star stencil and cond mode, separate or fused
refresh, plain lists and quad plans of 1..5 levels, fields at 16-byte
aligned offsets inside interleaved storages.  Bit-exact: FP32 FMAs in the
fixed tap order, round to nearest even on every store."""
import numpy as np
import pytest

import oracle as O
from bricklib_amd import BK_BF16, stencils as S

pytestmark = pytest.mark.gpu

NAMES = {0: "7pt", 1: "13pt", 2: "25pt", 4: "cond"}


def _cases(n=24, seed=2026):
    rng = np.random.default_rng(seed)
    for c in range(n):
        nb = tuple(int(x) for x in rng.integers(1, 6, size=3))
        kind = int(rng.choice([S.BK_STAR7, S.BK_STAR13, S.BK_STAR25, S.BK_COND7]))
        cmode = int(rng.integers(0, 2))
        plan = None if rng.random() < 0.4 else int(rng.integers(1, 6))
        fused = bool(rng.random() < 0.5)
        nsweeps = int(rng.integers(1, 4))
        off = int(rng.choice([0, 8, 256]))
        yield c, nb, kind, cmode, plan, fused, nsweeps, off


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: f"case{c[0]}")
def test_bf16_row_form_random_cases(gpu, case):
    c, nb, kind, cmode, plan, fused, nsweeps, off = case
    print(f"[case] bf16 rows #{c} nb={nb} {NAMES[kind]} cond_mode={cmode} quad={plan} "
          f"fused={fused} sweeps={nsweeps} offset={off}", flush=True)
    st = S.Stencil.make(kind, seed=100 + c, cond_mode=cmode)
    g = gpu.PeriodicGrid(nb)
    step = 1024  # two fields' room per chunk; this field at `off`
    sa = gpu.BrickStorage.allocate(g.nbricks, step, BK_BF16)
    sb = gpu.BrickStorage.allocate(g.nbricks, step, BK_BF16)
    sa.buf.zero()
    sb.buf.zero()
    a = gpu.Brick(g.info, sa, off)
    b = gpu.Brick(g.info, sb, off)
    g.fill_uniform(a, 300 + c)
    pl = None if plan is None else g.plan("quad", plan)
    res = gpu.sweeps(st, a, b, g, nsweeps, plan=pl, fused=fused)
    got = res.storage.to_host()[:, off:off + 512]
    og = O.Grid(nb)
    oa, ob = og.storage(BK_BF16), og.storage(BK_BF16)
    og.fill_uniform(oa, 300 + c, dtype=BK_BF16)
    want = og.sweeps(kind, st.coeff, oa, ob, nsweeps, dtype=BK_BF16, theta=st.theta,
                     cond_mode=st.cond_mode)
    np.testing.assert_array_equal(got[og.interior], want[og.interior])
    if fused:
        pr = g.pairs.reshape(-1, 2)
        np.testing.assert_array_equal(got[pr[:, 0]], got[pr[:, 1]])
