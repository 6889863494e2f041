"""GPU parity: the HIP sweeps (through the C-ABI) against the oracle.

Bar: bit-exact against the oracle's fixed-order (FMA) evaluation for every
stencil and dtype; FP64 within 1e-12 (one sweep) / 1e-10 (100 sweeps) of
the reference-order evaluation, relative error per point with the floor
max(|ref|, 1) -- the inputs are U[-1,1], so the floor is the input scale
and matches the absolute branch of compareBrick (brickcompare.h:35-39).
"""
import numpy as np
import pytest

import oracle as O
from bricklib_amd import BK_F32, BK_F64, stencils as S

pytestmark = pytest.mark.gpu

STENCILS = [S.BK_STAR7, S.BK_STAR13, S.BK_STAR25, S.BK_CUBE125, S.BK_COND7]
DTYPES = [0, 1, 2]  # F64, F32, BF16
NAMES = {0: "7pt", 1: "13pt", 2: "25pt", 3: "125pt", 4: "cond"}


def setup(bk, nb, dtype, seed, fields=1):
    g = bk.PeriodicGrid(nb)
    st = g.field_storage(dtype, fields)
    a = bk.Brick(g.info, st, 0)
    sb = g.field_storage(dtype, fields)
    b = bk.Brick(g.info, sb, 0)
    g.fill_uniform(a, seed)
    sb.buf.zero()
    og = O.Grid(nb)
    oa = og.storage(dtype, fields)
    og.fill_uniform(oa, seed, dtype=dtype)
    return g, a, b, og, oa


def host(brick):
    return brick.storage.to_host()


@pytest.mark.parametrize("dtype", DTYPES)
def test_fill_matches_oracle(gpu, dtype):
    g, a, b, og, oa = setup(gpu, (2, 3, 2), dtype, 17)
    np.testing.assert_array_equal(host(a)[og.interior], oa[og.interior])


@pytest.mark.parametrize("dtype", DTYPES)
def test_ghost_refresh_matches_oracle(gpu, dtype):
    g, a, b, og, oa = setup(gpu, (3, 2, 2), dtype, 5)
    g.ghost_refresh(a)
    og.refresh(oa, dtype)
    np.testing.assert_array_equal(host(a), oa)


@pytest.mark.parametrize("stencil", STENCILS, ids=lambda s: NAMES[s])
@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32", "bf16"])
def test_one_sweep_bit_exact(gpu, stencil, dtype):
    print(f"[case] one sweep {NAMES[stencil]} dtype={dtype} nb=(2,3,4)", flush=True)
    st = S.Stencil.make(stencil, seed=7)
    g, a, b, og, oa = setup(gpu, (2, 3, 4), dtype, 3)
    g.ghost_refresh(a)
    gpu.sweep(st, b, a, g)
    og.refresh(oa, dtype)
    ob = og.storage(dtype)
    og.sweep(stencil, st.coeff, oa, ob, dtype=dtype, theta=st.theta, cond_mode=st.cond_mode)
    got = host(b)
    np.testing.assert_array_equal(got[og.interior], ob[og.interior])


@pytest.mark.parametrize("stencil", STENCILS, ids=lambda s: NAMES[s])
def test_fp64_within_1e12_of_reference_order(gpu, stencil):
    st = S.Stencil.make(stencil, seed=2)
    g, a, b, og, oa = setup(gpu, (3, 3, 3), 0, 8)
    g.ghost_refresh(a)
    gpu.sweep(st, b, a, g)
    og.refresh(oa)
    ob = og.storage()
    og.sweep(stencil, st.coeff, oa, ob, mode=O.MODE_REF, theta=st.theta, cond_mode=st.cond_mode)
    err = O.field_compare(host(b), ob, og.interior, floor=1.0)
    print(f"[err] {NAMES[stencil]} 1 sweep vs reference order: {err}")
    assert err["max_rel"] <= 1e-12


@pytest.mark.parametrize("stencil", STENCILS, ids=lambda s: NAMES[s])
@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32", "bf16"])
@pytest.mark.parametrize("plan", [None, ("column", 4), ("column", 3), ("column", 8),
                                  ("quad", 2), ("quad", 8)],
                         ids=["list", "kstack4", "kstack3", "kmarch8", "quad2", "quad8"])
def test_one_sweep_bit_exact_plans(gpu, stencil, dtype, plan):
    """Every schedule gives the oracle's bits: per-brick list, k-stacks,
    k-column marching, 2x2 column quads (odd dims give partial quads)."""
    print(f"[case] plan={plan} {NAMES[stencil]} dtype={dtype} nb=(3,4,5)", flush=True)
    st = S.Stencil.make(stencil, seed=11)
    g, a, b, og, oa = setup(gpu, (3, 4, 5), dtype, 13)
    g.ghost_refresh(a)
    pl = None if plan is None else g.plan(plan[0], kcol=plan[1])
    gpu.sweep(st, b, a, g, plan=pl)
    og.refresh(oa, dtype)
    ob = og.storage(dtype)
    og.sweep(stencil, st.coeff, oa, ob, dtype=dtype, theta=st.theta, cond_mode=st.cond_mode)
    np.testing.assert_array_equal(host(b)[og.interior], ob[og.interior])
    # bricks outside the interior list are never written
    ghosts = np.setdiff1d(np.arange(og.nbricks), og.interior)
    assert not host(b)[ghosts].any()


@pytest.mark.parametrize("shape", ["column", "quad"])
def test_plan_of_shuffled_list(gpu, shape):
    """A plan built from a shuffled interior list (broken chains, scattered
    quads) still writes exactly the oracle's values."""
    nb = (4, 4, 4)
    stencil = S.BK_CUBE125 if shape == "quad" else S.BK_STAR25
    st = S.Stencil.make(stencil, seed=2)
    g, a, b, og, oa = setup(gpu, nb, 0, 5)
    g.ghost_refresh(a)
    blist = g.interior.copy()
    np.random.default_rng(1).shuffle(blist)
    pl = gpu.Plan(g.info, blist, shape, 4)
    gpu.sweep(st, b, a, g, plan=pl)
    og.refresh(oa)
    ob = og.storage()
    og.sweep(stencil, st.coeff, oa, ob)
    np.testing.assert_array_equal(host(b)[og.interior], ob[og.interior])


@pytest.mark.parametrize("stencil", STENCILS, ids=lambda s: NAMES[s])
@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
def test_pingpong_sweeps_bit_exact(gpu, stencil, use_graph):
    print(f"[case] 8 sweeps {NAMES[stencil]} graph={use_graph}", flush=True)
    st = S.Stencil.make(stencil, seed=4)
    g, a, b, og, oa = setup(gpu, (2, 2, 3), 0, 12)
    res = gpu.sweeps(st, a, b, g, 8, use_graph=use_graph)
    ob = og.storage()
    ores = og.sweeps(stencil, st.coeff, oa, ob, 8, theta=st.theta, cond_mode=st.cond_mode)
    np.testing.assert_array_equal(host(res)[og.interior], ores[og.interior])


def test_graph_replays_identical(gpu):
    st = S.Stencil.make("7pt", seed=1)
    g, a, b, og, oa = setup(gpu, (2, 2, 2), 0, 1)
    ref = None
    for rep in range(4):
        g.fill_uniform(a, 1)
        gpu.sweeps(st, a, b, g, 4, use_graph=True)
        gpu.device_sync()
        cur = host(a)[og.interior]  # 4 sweeps: result back in a
        if ref is None:
            ref = cur
            g.fill_uniform(a, 1)
            gpu.sweeps(st, a, b, g, 4, use_graph=False)
            np.testing.assert_array_equal(host(a)[og.interior], ref)
        else:
            np.testing.assert_array_equal(cur, ref)


@pytest.mark.parametrize("stencil", [S.BK_STAR7, S.BK_CUBE125], ids=["7pt", "125pt"])
def test_interleaved_fields(gpu, stencil):
    """bIn at offset 0 and bOut at offset 512 of one storage with step 1024,
    as stencils/3axis.cpp:26-28 lays them out."""
    st = S.Stencil.make(stencil, seed=9)
    g = gpu.PeriodicGrid((2, 2, 2))
    storage = g.field_storage(0, fields=2)
    storage.buf.zero()
    bin_, bout = gpu.Brick(g.info, storage, 0), gpu.Brick(g.info, storage, 512)
    g.fill_uniform(bin_, 31)
    g.ghost_refresh(bin_)
    gpu.sweep(st, bout, bin_, g)
    og = O.Grid((2, 2, 2))
    oa = og.storage()
    og.fill_uniform(oa, 31)
    og.refresh(oa)
    ob = og.storage()
    og.sweep(stencil, st.coeff, oa, ob)
    h = storage.to_host()
    np.testing.assert_array_equal(h[og.interior, 512:], ob[og.interior])
    np.testing.assert_array_equal(h[og.interior, :512], oa[og.interior])


@pytest.mark.parametrize("stencil", [S.BK_STAR7, S.BK_STAR13, S.BK_STAR25, S.BK_COND7],
                         ids=lambda s: NAMES[s])
def test_bf16_row_form_and_unaligned_fields(gpu, stencil):
    """BF16 sweeps -- per brick, and in 2x2 blocks of 1 and 3 levels --
    move each i-row as one 16-byte load and store when the fields and the
    brick step keep rows 16-byte aligned, and fall back to the per-cell
    kernels otherwise (a step that is not a multiple of 8 elements, a field
    at an odd offset): the oracle's bits either way, and the other field of
    the storage untouched."""
    from bricklib_amd import BK_BF16
    st = S.Stencil.make(stencil, seed=29)
    nb = (3, 2, 3)
    og = O.Grid(nb)
    oa = og.storage(BK_BF16)
    og.fill_uniform(oa, 37, dtype=BK_BF16)
    og.refresh(oa, BK_BF16)
    ob = og.storage(BK_BF16)
    og.sweep(stencil, st.coeff, oa, ob, dtype=BK_BF16, theta=st.theta, cond_mode=st.cond_mode)
    g = gpu.PeriodicGrid(nb)
    for step, off_in, off_out in ((1024, 0, 512), (1028, 0, 514), (1040, 4, 520),
                                  (1032, 520, 8)):
        for plan in (None, 1, 3):
            print(f"[case] bf16 rows {NAMES[stencil]} step={step} in@{off_in} out@{off_out} "
                  f"quad levels={plan}", flush=True)
            storage = gpu.BrickStorage.allocate(g.nbricks, step, BK_BF16)
            storage.buf.zero()
            bin_ = gpu.Brick(g.info, storage, off_in)
            bout = gpu.Brick(g.info, storage, off_out)
            g.fill_uniform(bin_, 37)
            g.ghost_refresh(bin_)
            gpu.sweep(st, bout, bin_, g, plan=None if plan is None else g.plan("quad", plan))
            h = storage.to_host()
            np.testing.assert_array_equal(h[og.interior, off_out:off_out + 512],
                                          ob[og.interior])
            np.testing.assert_array_equal(h[:, off_in:off_in + 512], oa)


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32", "bf16"])
def test_empty_brick_list_and_plan_are_no_ops(gpu, dtype):
    """An empty brick list, and a plan built from one, sweep nothing: the
    call succeeds, launches no kernel with an empty grid and leaves the
    output untouched (per-brick, quad and column schedules, every stencil)."""
    import ctypes as C
    from bricklib_amd import _lib
    from bricklib_amd.bricks import _desc
    g, a, b, og, oa = setup(gpu, (2, 2, 3), dtype, 21)
    empty = np.empty(0, np.uint32)
    for stencil in STENCILS:
        st = S.Stencil.make(stencil, seed=3)
        for plan in (None, gpu.Plan(g.info, empty, "quad", 2),
                     gpu.Plan(g.info, empty, "column", 4)):
            print(f"[case] empty list {NAMES[stencil]} dtype={dtype} "
                  f"plan={None if plan is None else plan.shape}", flush=True)
            keep = []
            d = _desc(st, b, a, g, keep, plan=plan)
            if plan is None:
                d.nlist = 0
            _lib.check(_lib.load().bk_sweep(C.byref(d), None))
            assert not host(b).any()


@pytest.mark.parametrize("nb", [(1, 1, 1), (1, 2, 5), (7, 3, 1)])
def test_ragged_and_single_brick_grids(gpu, nb):
    """Grids with one-brick axes (a brick that is its own neighbour, partial
    2x2 blocks and levels) under the default schedules of every stencil, in
    FP64 and in BF16 (whose defaults are the row-form kernels)."""
    from bricklib_amd import BK_BF16
    for dtype in (BK_F64, BK_BF16):
        for stencil in STENCILS:
            print(f"[case] ragged nb={nb} {NAMES[stencil]} dtype={dtype}", flush=True)
            st = S.Stencil.make(stencil, seed=6)
            g, a, b, og, oa = setup(gpu, nb, dtype, 2)
            res = gpu.sweeps(st, a, b, g, 3)
            ob = og.storage(dtype)
            ores = og.sweeps(stencil, st.coeff, oa, ob, 3, dtype=dtype, theta=st.theta,
                             cond_mode=st.cond_mode)
            np.testing.assert_array_equal(host(res)[og.interior], ores[og.interior])


def test_plain_case_128(gpu):
    """Plain method at 128^3: 7-point star, FP64, coefficients U(0,1)
    normalised, inputs U[-1,1], periodic ghost bricks; 1 sweep, then 8
    ping-pong sweeps; against the oracle's brick loop (bit-exact) and its
    array loop in the reference order (<= 1e-12)."""
    nb = (16, 16, 16)
    coeff = S.coeff_plain7(2024)
    st = S.Stencil(S.BK_STAR7, coeff)
    g, a, b, og, oa = setup(gpu, nb, 0, 99)
    g.ghost_refresh(a)
    gpu.sweep(st, b, a, g)
    og.refresh(oa)
    ob = og.storage()
    og.sweep(0, coeff, oa, ob)
    np.testing.assert_array_equal(host(b)[og.interior], ob[og.interior])
    obr = og.storage()
    og.sweep(0, coeff, oa, obr, mode=O.MODE_REF)
    e1 = O.field_compare(host(b), obr, og.interior, floor=1.0)
    assert e1["max_rel"] <= 1e-12
    # 8 ping-pong sweeps from the same input
    g.fill_uniform(a, 99)
    res = gpu.sweeps(st, a, b, g, 8)
    oa2 = og.storage()
    og.fill_uniform(oa2, 99)
    ob2 = og.storage()
    ores = og.sweeps(0, coeff, oa2, ob2, 8)
    np.testing.assert_array_equal(host(res)[og.interior], ores[og.interior])
    oa3 = og.storage()
    og.fill_uniform(oa3, 99)
    ob3 = og.storage()
    oref = og.sweeps(0, coeff, oa3, ob3, 8, mode=O.MODE_REF)
    e8 = O.field_compare(host(res), oref, og.interior, floor=1.0)
    print(f"[err] plain 128^3: 1 sweep {e1['max_rel']:.3e}, 8 sweeps {e8['max_rel']:.3e}")
    assert e8["max_rel"] <= 1e-10


def test_device_reductions(gpu):
    g, a, b, og, oa = setup(gpu, (3, 2, 2), 0, 42)
    s = g.stats(a)
    x = oa[og.interior].reshape(-1)
    assert abs(s[0] - x.sum()) < 1e-9
    assert abs(s[1] - (x * x).sum()) < 1e-9
    assert s[2] == np.abs(x).max()
    g2 = g.stats(a)
    np.testing.assert_array_equal(s, g2)  # deterministic
    c = g.compare(a, a)
    assert c["max_abs"] == 0 and c["bad"] == 0


def test_device_copy_to_from_brick(gpu):
    dims, pad, ghost = [16, 24, 8], [8, 8, 8], [8, 8, 8]
    bdims = [(d + 2 * gh) // 8 for d, gh in zip(dims, ghost)]
    grid, info = gpu.init_grid(bdims)
    ext = [d + 2 * (p + gh) for d, p, gh in zip(dims, pad, ghost)]
    arr = np.random.default_rng(0).standard_normal(int(np.prod(ext)))
    d_arr = gpu.DeviceBuffer.from_array(arr)
    d_grid = gpu.DeviceBuffer.from_array(grid.reshape(-1))
    st = gpu.BrickStorage.allocate(info.nbricks, 512)
    st.buf.zero()
    br = gpu.Brick(info, st, 0)
    gpu.copyToBrick(dims, pad, ghost, d_arr, d_grid, br)
    exp = np.zeros((info.nbricks, 512))
    O.lib().or_copy_to_brick(0, O.L3(dims), O.L3(pad), O.L3(ghost), O.P(arr),
                             O.P(grid.reshape(-1), O.u32p), O.P(exp), 512)
    np.testing.assert_array_equal(st.to_host(), exp)
    out = gpu.DeviceBuffer(arr.nbytes)
    out.zero()
    gpu.copyFromBrick(dims, pad, ghost, out, d_grid, br)
    back = out.download(np.float64)
    expb = np.zeros_like(arr)
    O.lib().or_copy_from_brick(0, O.L3(dims), O.L3(pad), O.L3(ghost), O.P(expb),
                               O.P(grid.reshape(-1), O.u32p), O.P(exp), 512)
    np.testing.assert_array_equal(back, expb)


def test_aliasing_rejected(gpu):
    st = S.Stencil.make("7pt")
    g, a, b, og, oa = setup(gpu, (1, 1, 1), 0, 1)
    with pytest.raises(gpu.BrickError):
        gpu.sweep(st, a, a, g)


def test_stream_copy(gpu):
    import ctypes as C
    from bricklib_amd import _lib
    lib = gpu.load()
    for nbytes in (16, 4096 * 3 + 48, 1 << 20):
        src = np.random.default_rng(nbytes).integers(0, 255, nbytes, dtype=np.uint8)
        ds = gpu.DeviceBuffer.from_array(src)
        dd = gpu.DeviceBuffer(nbytes)
        dd.zero()
        _lib.check(lib.bk_stream_copy(C.c_void_p(dd.ptr), C.c_void_p(ds.ptr), nbytes, None))
        np.testing.assert_array_equal(dd.download(np.uint8), src)
    assert lib.bk_stream_copy(C.c_void_p(dd.ptr), C.c_void_p(dd.ptr), 16, None) == 1


@pytest.mark.parametrize("name", ["7pt", "25pt", "125pt"])
def test_sweep_order_does_not_change_bits(gpu, name):
    """Grid order, square and rectangular tiles visit the interior bricks in
    different orders (each a permutation of the interior); the sweeps give
    the same bits."""
    st = S.Stencil.make(name, seed=12)
    outs = []
    for tile in (0, 3, (4, 2), [2, 5]):
        g = gpu.PeriodicGrid((7, 5, 6), tile=tile)
        assert np.array_equal(np.sort(g.sweep_list), np.sort(g.interior))
        a = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
        b = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
        g.fill_uniform(a, 21)
        outs.append(gpu.sweeps(st, a, b, g, 3).storage.to_host()[g.interior])
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


@pytest.mark.parametrize("dtype,chains", [(BK_F64, 1), (BK_F64, 8), (BK_F32, 1), (BK_F32, 8)])
def test_valu_probe(gpu, dtype, chains):
    import ctypes as C
    from bricklib_amd import _lib
    lib = gpu.load()
    n = 256 * 4 * 64  # one wave per SIMD on up to 256 CUs
    npt = np.float64 if dtype == BK_F64 else np.float32
    out = gpu.DeviceBuffer(n * np.dtype(npt).itemsize)
    fl = C.c_double(0)
    iters = 3
    _lib.check(lib.bk_valu_probe(dtype, chains, 1, iters, C.c_void_p(out.ptr), n, C.byref(fl),
                                 None))
    gpu.device_sync()
    nblocks = round(fl.value / (2 * 64 * iters * 64))
    assert fl.value == 2.0 * 64 * iters * nblocks * 64 and nblocks % 4 == 0
    got = out.download(npt)[:nblocks * 64].astype(np.float64).reshape(nblocks, 64)
    # every block is the same: lane t sums `chains` accumulators
    # acc_i = t/1000 + i + iters * (32/chains) * x_i * (0.999 + 1.0001), x_i = (t + i + 1)/2
    t = np.arange(64, dtype=np.float64)[:, None]
    i = np.arange(chains, dtype=np.float64)[None, :]
    want = (t / 1000 + i + iters * (32 / chains) * (t + i + 1) / 2 * (0.999 + 1.0001)).sum(axis=1)
    np.testing.assert_allclose(got, np.broadcast_to(want, got.shape),
                               rtol=1e-12 if dtype == BK_F64 else 1e-4)
    # error paths: too small an output, bad dtype or chain count
    assert lib.bk_valu_probe(dtype, chains, 1, iters, C.c_void_p(out.ptr), 64, C.byref(fl),
                             None) == 1
    assert lib.bk_valu_probe(2, chains, 1, iters, C.c_void_p(out.ptr), n, C.byref(fl), None) == 1
    assert lib.bk_valu_probe(dtype, 3, 1, iters, C.c_void_p(out.ptr), n, C.byref(fl), None) == 1


DEEP = [("7pt", 8), ("13pt", 4), ("25pt", 2), ("125pt", 4), ("cond", 8)]


@pytest.mark.parametrize("name,gi", DEEP)
@pytest.mark.parametrize("use_graph", [False, True])
def test_deep_ghost_matches_refresh_every_sweep(gpu, name, gi, use_graph):
    """bk_sweeps_deep (refresh once per ghost_iters sweeps, sweeps over every
    brick in between) gives the refresh-every-sweep result bit for bit, for
    the largest ghost_iters the 8-cell ghost zone allows and a sweep count
    that is not a multiple of it."""
    from bricklib_amd import BK_F32
    nb = (3, 2, 4)
    st = S.Stencil.make(name, seed=21)
    n = 2 * gi + 3
    for dt in ((BK_F64, BK_F32) if name in ("7pt", "125pt") else (BK_F64,)):
        g = gpu.PeriodicGrid(nb)
        outs = []
        for k in (1, gi):
            a = gpu.Brick(g.info, g.field_storage(dt), 0)
            b = gpu.Brick(g.info, g.field_storage(dt), 0)
            g.fill_uniform(a, 5)
            res = gpu.sweeps(st, a, b, g, n, use_graph=use_graph, ghost_iters=k)
            outs.append(res.storage.to_host()[g.interior])
        print(f"[case] deep {name} gi={gi} dt={dt} graph={use_graph}", flush=True)
        np.testing.assert_array_equal(outs[1], outs[0])


def test_deep_ghost_rejects_too_many_sweeps(gpu):
    from bricklib_amd import BrickError
    g = gpu.PeriodicGrid((2, 2, 2))
    st = S.Stencil.make("25pt", seed=2)
    a = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    b = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    with pytest.raises(BrickError):
        gpu.sweeps(st, a, b, g, 4, ghost_iters=3)


@pytest.mark.parametrize("name", ["7pt", "13pt", "25pt", "cond"])
@pytest.mark.parametrize("use_graph", [False, True])
def test_fused_refresh_matches_refresh_every_sweep(gpu, name, use_graph):
    """bk_sweeps_fused (ghosts refreshed once, then written by each sweep's
    stores) gives the refresh-every-sweep result bit for bit, for odd and
    even sweep counts, on grids down to one brick (26 ghost images)."""
    from bricklib_amd import BK_BF16, BK_F32
    st = S.Stencil.make(name, seed=31)
    for nb in ((3, 2, 4), (1, 1, 1), (2, 5, 1)):
        for dt in (BK_F64, BK_F32, BK_BF16):
            g = gpu.PeriodicGrid(nb)
            for n in (1, 4, 5):
                outs = []
                for fused in (False, True):
                    a = gpu.Brick(g.info, g.field_storage(dt), 0)
                    b = gpu.Brick(g.info, g.field_storage(dt), 0)
                    g.fill_uniform(a, 6)
                    res = gpu.sweeps(st, a, b, g, n, use_graph=use_graph, plan=None,
                                     fused=fused)
                    outs.append(res.storage.to_host())
                print(f"[case] fused {name} nb={nb} dt={dt} n={n} graph={use_graph}", flush=True)
                np.testing.assert_array_equal(outs[1][g.interior], outs[0][g.interior])
                # the fused form leaves the result's ghosts refreshed too
                pr = g.pairs.reshape(-1, 2)
                np.testing.assert_array_equal(outs[1][pr[:, 0]], outs[1][pr[:, 1]])


@pytest.mark.parametrize("name", ["7pt", "13pt", "25pt", "cond"])
@pytest.mark.parametrize("kcol", [1, 3, 8])
def test_fused_refresh_quads_match_oracle(gpu, orc, name, kcol):
    """The fused refresh in the quad-column star kernel: the result of 1, 4
    and 5 sweeps is the oracle's (refresh every sweep), bit for bit, on odd
    grids (partial quads and levels), and the result's ghosts are its
    periodic images."""
    from bricklib_amd import BK_BF16, BK_F32
    st = S.Stencil.make(name, seed=33)
    for nb in ((3, 2, 5), (2, 2, 2)):
        g = gpu.PeriodicGrid(nb)
        og = O.Grid(nb)
        for dt in (BK_F64, BK_F32, BK_BF16):
            for n in (1, 4, 5):
                print(f"[case] fused quads {name} kcol={kcol} nb={nb} dt={dt} n={n}", flush=True)
                a = gpu.Brick(g.info, g.field_storage(dt), 0)
                b = gpu.Brick(g.info, g.field_storage(dt), 0)
                g.fill_uniform(a, 8)
                res = gpu.sweeps(st, a, b, g, n, use_graph=(n == 4), plan=g.plan("quad", kcol),
                                 fused=True)
                got = res.storage.to_host()
                oa, ob = og.storage(dt), og.storage(dt)
                og.fill_uniform(oa, 8, dtype=dt)
                want = og.sweeps(st.kind, st.coeff, oa, ob, n, dtype=dt, theta=st.theta,
                                 cond_mode=st.cond_mode)
                np.testing.assert_array_equal(got[og.interior], want[og.interior])
                pr = g.pairs.reshape(-1, 2)
                np.testing.assert_array_equal(got[pr[:, 0]], got[pr[:, 1]])


@pytest.mark.parametrize("stencil", [S.BK_STAR7, S.BK_STAR13, S.BK_STAR25, S.BK_COND7],
                         ids=lambda s: NAMES[s])
@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32", "bf16"])
@pytest.mark.parametrize("kcol", [1, 3, 8])
def test_lane_shift_quads_bit_exact(gpu, stencil, dtype, kcol):
    """The quad kernels with the i-taps from DPP lane shifts (lane_shift=1)
    give the oracle's bits, on grids whose extents leave partial quads and
    levels, including single-brick axes (a member that is its own
    i-neighbour); bricks outside the interior are never written."""
    st = S.Stencil.make(stencil, seed=19)
    for nb in ((3, 4, 5), (1, 3, 2), (2, 1, 4)):
        print(f"[case] lane_shift {NAMES[stencil]} dtype={dtype} kcol={kcol} nb={nb}",
              flush=True)
        g, a, b, og, oa = setup(gpu, nb, dtype, 23)
        g.ghost_refresh(a)
        gpu.sweep(st, b, a, g, plan=g.plan("quad", kcol), lane_shift=1)
        og.refresh(oa, dtype)
        ob = og.storage(dtype)
        og.sweep(stencil, st.coeff, oa, ob, dtype=dtype, theta=st.theta, cond_mode=st.cond_mode)
        np.testing.assert_array_equal(host(b)[og.interior], ob[og.interior])
        ghosts = np.setdiff1d(np.arange(og.nbricks), og.interior)
        assert not host(b)[ghosts].any()


@pytest.mark.parametrize("name", ["7pt", "25pt", "cond"])
def test_lane_shift_fused_sweeps_match_oracle(gpu, name):
    """Several sweeps with lane shifts and the fused refresh, eagerly and
    replayed as a graph: the oracle's result, bit for bit."""
    st = S.Stencil.make(name, seed=41)
    nb = (3, 2, 5)
    g = gpu.PeriodicGrid(nb)
    og = O.Grid(nb)
    for dt in (BK_F64, BK_F32):
        for n, graph in ((1, False), (5, False), (4, True)):
            print(f"[case] lane_shift fused {name} dt={dt} n={n} graph={graph}", flush=True)
            a = gpu.Brick(g.info, g.field_storage(dt), 0)
            b = gpu.Brick(g.info, g.field_storage(dt), 0)
            g.fill_uniform(a, 6)
            res = gpu.sweeps(st, a, b, g, n, use_graph=graph, plan=g.plan("quad", 2), fused=True,
                             lane_shift=1)
            got = res.storage.to_host()
            oa, ob = og.storage(dt), og.storage(dt)
            og.fill_uniform(oa, 6, dtype=dt)
            want = og.sweeps(st.kind, st.coeff, oa, ob, n, dtype=dt, theta=st.theta,
                             cond_mode=st.cond_mode)
            np.testing.assert_array_equal(got[og.interior], want[og.interior])


def test_lane_shift_rejected_outside_quad_stars(gpu):
    from bricklib_amd import BrickError
    g = gpu.PeriodicGrid((2, 2, 2))
    a = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    b = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    with pytest.raises(BrickError):
        gpu.sweep(S.Stencil.make("125pt", seed=1), b, a, g, plan=g.plan("quad", 2), lane_shift=1)
    with pytest.raises(BrickError):
        gpu.sweep(S.Stencil.make("7pt", seed=1), b, a, g, plan=g.plan("column", 4), lane_shift=1)
    with pytest.raises(BrickError):
        gpu.sweep(S.Stencil.make("7pt", seed=1), b, a, g, plan=None, lane_shift=1)


def test_fused_refresh_rejects_plans_and_cube(gpu):
    from bricklib_amd import BrickError
    g = gpu.PeriodicGrid((2, 2, 2))
    a = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    b = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    with pytest.raises(BrickError):
        gpu.sweeps(S.Stencil.make("125pt", seed=1), a, b, g, 2, fused=True)
    with pytest.raises(BrickError):
        gpu.sweeps(S.Stencil.make("25pt", seed=1), a, b, g, 2, fused=True,
                   plan=g.plan("column", 8))


@pytest.mark.parametrize("name", ["13pt", "125pt"])
def test_tuner_picks_a_schedule_with_the_same_bits(gpu, name):
    g = gpu.PeriodicGrid((4, 4, 12))
    st = S.Stencil.make(name, seed=8)
    a = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    b = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    g.fill_uniform(a, 2)
    before = gpu.sweeps(st, a, b, g, 3).storage.to_host()[g.interior]
    out = g.tune(st, BK_F64, steps=2, warm_s=0.05)
    if name == "125pt":
        want = {"8x8x8 brick wg4", "8x8x8 quad2", "8x8x8 quad4", "8x8x8 quad6", "8x8x8 quad8"}
    else:
        want = {"8x8x8 brick wg1", "8x8x8 brick wg2", "8x8x8 brick wg4", "8x8x8 brick wg8",
                "8x8x8 column4", "8x8x8 column8", "8x8x8 column16", "8x8x8 quad1",
                "8x8x8 quad2", "8x8x8 quad4", "8x8x8 quad8", "8x8x8 quad1+dpp",
                "8x8x8 quad2+dpp", "8x8x8 quad4+dpp", "8x8x8 quad8+dpp"}
    assert set(out["ms"]) == want
    assert out["best"] in out["ms"]
    g.fill_uniform(a, 2)
    after = gpu.sweeps(st, a, b, g, 3).storage.to_host()[g.interior]
    np.testing.assert_array_equal(after, before)
