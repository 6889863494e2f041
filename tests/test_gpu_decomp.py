"""Sweeps on the BrickDecomp layout (refresh = the 26 neighbours' ghost runs
copied from their skin runs) give the init_grid layout's values bit for
bit, through every schedule and form."""
import numpy as np
import pytest

from bricklib_amd import BK_F32, BK_F64, stencils as S
from bricklib_amd.decomp import SKIN3D_BAD, SKIN3D_NORMAL, DecompGrid

pytestmark = pytest.mark.gpu


def run(gpu, grid, st, dt, n, **kw):
    a = gpu.Brick(grid.info, grid.field_storage(dt), 0)
    b = gpu.Brick(grid.info, grid.field_storage(dt), 0)
    grid.fill_uniform(a, 17)
    res = gpu.sweeps(st, a, b, grid, n, **kw)
    return res.storage.to_host()[grid.interior]


@pytest.mark.parametrize("name", ["7pt", "13pt", "25pt", "125pt", "cond"])
def test_decomp_matches_init_grid_layout(gpu, name):
    nb = (4, 3, 5)
    st = S.Stencil.make(name, seed=41)
    dg = DecompGrid(nb)
    pg = gpu.PeriodicGrid(nb, tile=0)
    for dt in ((BK_F64, BK_F32) if name in ("7pt", "125pt") else (BK_F64,)):
        print(f"[case] decomp {name} dt={dt}", flush=True)
        want = run(gpu, pg, st, dt, 3)
        np.testing.assert_array_equal(run(gpu, dg, st, dt, 3), want)
        if name != "125pt":
            np.testing.assert_array_equal(run(gpu, dg, st, dt, 3, plan=None, fused=True), want)
        gi = 8 // S.RADIUS[st.kind]
        np.testing.assert_array_equal(run(gpu, dg, st, dt, 3, ghost_iters=min(gi, 3)), want)


@pytest.mark.parametrize("skin", ["normal", "bad"])
def test_other_skin_orders_match(gpu, skin):
    """The reference's other surface orders number the bricks differently;
    the sweeps over them give the same values bit for bit."""
    nb = (3, 4, 5)
    st = S.Stencil.make("13pt", seed=7)
    dg = DecompGrid(nb, skin={"normal": SKIN3D_NORMAL, "bad": SKIN3D_BAD}[skin])
    pg = gpu.PeriodicGrid(nb, tile=0)
    print(f"[case] decomp skin={skin}", flush=True)
    np.testing.assert_array_equal(run(gpu, dg, st, BK_F64, 4), run(gpu, pg, st, BK_F64, 4))
