"""GPU sweeps against the golden fixtures: every stencil, both grids, 1 and
8 sweeps, the FMA-order digests bit for bit -- through the default schedule
of the ghost-zone layout, the per-brick schedule, and the wrapped-adjacency
layout."""
import json
import os

import numpy as np
import pytest

from bricklib_amd import BK_F64, stencils as S
from bricklib_amd.batch import Request  # noqa: F401  (package import check)

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(HERE, "sweeps.json")))
ARR = np.load(os.path.join(HERE, "sweeps.npz"))


def dense(bricks: np.ndarray, nb) -> np.ndarray:
    """Interior bricks in grid order -> dense (z, y, x) field."""
    nb0, nb1, nb2 = nb
    return bricks.reshape(nb2, nb1, nb0, 8, 8, 8).transpose(0, 3, 1, 4, 2, 5).reshape(
        nb2 * 8, nb1 * 8, nb0 * 8)


def digest(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, "<f8").tobytes()).hexdigest()


@pytest.mark.parametrize("cname", sorted(META["cases"]))
@pytest.mark.parametrize("layout", ["ghost_auto", "ghost_perbrick", "wrapped"])
def test_gpu_matches_golden(gpu, cname, layout):
    case = META["cases"][cname]
    st = S.Stencil(kind=case["kind"], coeff=np.array(ARR[f"coeff_{cname}"]),
                   theta=float(ARR[f"theta_{cname}"][0]), cond_mode=case["cond_mode"])
    for gname, nb in META["grids"].items():
        g = gpu.PeriodicGrid(nb, ghost=0 if layout == "wrapped" else 1)
        for ns in (1, 8):
            a = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
            b = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
            g.fill_uniform(a, META["seed_in"])
            plan = None if layout == "ghost_perbrick" else "auto"
            res = gpu.sweeps(st, a, b, g, ns, plan=plan)
            out = dense(res.storage.to_host()[g.interior], nb)
            print(f"[case] {cname} {gname} sweeps={ns} {layout}", flush=True)
            if gname == "16" and ns == 1:
                np.testing.assert_array_equal(out, ARR[f"out_{cname}_16_1_fma"])
            assert digest(out) == case["out"][f"{gname}/{ns}/fma"]["sha256"]
