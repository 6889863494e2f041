"""The oracle against what the mathematics fixes.

- brute force: every stencil on a dense periodic field with numpy.roll
  (a textbook periodic stencil, independent of bricks and adjacency);
- closed forms: constant fields, plane waves (eigenvectors of symmetric
  stars with eigenvalue c0 + sum 2 c_r cos(2 pi m r / n) per axis), cond on
  all-negative / all-positive fields;
- the array loop (3axis.cpp:32-37) equals the brick loop;
- ghost refresh leaves ghosts equal to their periodic image;
- planted mistakes (in the oracle's inputs, on the CPU) are caught.
"""
import numpy as np
import pytest

import oracle as O
from bricklib_amd import stencils as S


def dense_ref(field, stencil, coeff, theta=0.0, cond_mode=0):
    """numpy brute force on a periodic (z, y, x) field."""
    offs = S.offsets(stencil)
    relu = stencil == O.COND7
    acc = np.zeros_like(field)
    for c, (di, dj, dk) in zip(coeff, offs):
        x = np.roll(field, shift=(-dk, -dj, -di), axis=(0, 1, 2))
        if relu:
            x = np.where(x < 0, 0.0, x)
        acc += c * x
    if relu:
        sel = acc > 0 if cond_mode == 0 else field > theta
        acc = np.where(sel, acc, -acc)
    return acc


STENCILS = [O.STAR7, O.STAR13, O.STAR25, O.CUBE125, O.COND7]


def make(g, seed=3):
    a = g.storage()
    g.fill_uniform(a, seed)
    return a


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("mode", [O.MODE_FMA, O.MODE_REF])
def test_one_sweep_vs_numpy_bruteforce(stencil, mode):
    g = O.Grid((2, 3, 4))
    a = make(g)
    coeff = np.random.default_rng(stencil).uniform(-1, 1, O.lib().or_stencil_ncoeff(stencil))
    theta = 0.1
    b = g.storage()
    g.refresh(a)
    g.sweep(stencil, coeff, a, b, mode=mode, theta=theta, cond_mode=1)
    got = g.to_dense(b)
    exp = dense_ref(g.to_dense(a), stencil, coeff, theta, 1)
    np.testing.assert_allclose(got, exp, rtol=0, atol=1e-13)


@pytest.mark.parametrize("stencil", [O.STAR7, O.STAR25, O.CUBE125])
def test_eight_pingpong_sweeps_vs_numpy(stencil):
    g = O.Grid((3, 2, 2))
    a = make(g, 9)
    coeff = S.default_coeff(stencil, 4)
    f = g.to_dense(a)
    b = g.storage()
    res = g.sweeps(stencil, coeff, a, b, 8)
    for _ in range(8):
        f = dense_ref(f, stencil, coeff)
    np.testing.assert_allclose(g.to_dense(res), f, rtol=0, atol=1e-13)


@pytest.mark.parametrize("stencil", STENCILS)
def test_constant_field(stencil):
    g = O.Grid((2, 2, 2))
    a = g.storage()
    a[:] = 0.75
    coeff = S.default_coeff(stencil, 2)
    b = g.storage()
    g.sweep(stencil, coeff, a, b, cond_mode=0)
    exp = 0.75 * coeff.sum()
    assert np.allclose(g.to_dense(b), exp, rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("stencil", [O.STAR7, O.STAR13, O.STAR25])
def test_plane_wave_eigenvalue_100_sweeps(stencil):
    """cos(2 pi (m.x)/n) is an eigenvector of a per-axis-symmetric star."""
    nb = (2, 2, 3)
    g = O.Grid(nb)
    n = [8 * x for x in nb]
    m = (1, 2, 1)
    z, y, x = np.meshgrid(*[np.arange(v) for v in n[::-1]], indexing="ij")
    f = np.cos(2 * np.pi * (m[0] * x / n[0] + m[1] * y / n[1] + m[2] * z / n[2]))
    coeff = S.coeff_star(stencil, 5)
    R = S.RADIUS[stencil]
    lam = coeff[0]
    for r in range(1, R + 1):
        base = 6 * (r - 1) + 1
        for ax in range(3):
            lam += 2 * coeff[base + 2 * ax] * np.cos(2 * np.pi * m[ax] * r / n[ax])
    a = g.storage()
    g.from_dense(f, a)
    b = g.storage()
    res = g.sweeps(stencil, coeff, a, b, 100)
    np.testing.assert_allclose(g.to_dense(res), lam ** 100 * f, rtol=0, atol=1e-12)


def test_cond_closed_forms():
    g = O.Grid((2, 2, 2))
    coeff = S.coeff_plain7(1)
    a = g.storage()
    g.fill_uniform(a, 5, lo=-1.0, hi=-0.1)  # all negative: relu -> 0
    g.refresh(a)
    b = g.storage()
    g.sweep(O.COND7, coeff, a, b, cond_mode=0)
    assert np.all(g.to_dense(b) == 0.0)
    g.fill_uniform(a, 5, lo=0.1, hi=1.0)  # all positive: |7pt|
    g.refresh(a)
    c = g.storage()
    g.sweep(O.COND7, coeff, a, b, cond_mode=0)
    g.sweep(O.STAR7, coeff, a, c)
    np.testing.assert_array_equal(g.to_dense(b), np.abs(g.to_dense(c)))


@pytest.mark.parametrize("stencil", STENCILS)
def test_array_loop_equals_brick_loop(stencil):
    """The dense array loop (3axis.cpp:32-37) with a periodic 8-cell ghost
    zone gives bit-identical results to the brick loop (same order)."""
    nb = (2, 3, 2)
    g = O.Grid(nb)
    a = make(g, 21)
    g.refresh(a)
    coeff = S.default_coeff(stencil, 8)
    b = g.storage()
    g.sweep(stencil, coeff, a, b, mode=O.MODE_REF, theta=0.05, cond_mode=1)
    n = [8 * v for v in nb]
    G = 8
    ext = [v + 2 * G for v in n]
    arr = np.zeros(ext[::-1])
    arr[G:G + n[2], G:G + n[1], G:G + n[0]] = g.to_dense(a)
    O.lib().or_array_periodic_fill(O.P(arr, O.dblp), O.L3(n), G)
    out = np.zeros_like(arr)
    c = np.ascontiguousarray(coeff)
    O.lib().or_array_sweep(stencil, O.MODE_REF, O.L3(n), G, O.P(arr, O.dblp), O.P(out, O.dblp),
                           O.P(c, O.dblp), 0.05, 1)
    np.testing.assert_array_equal(out[G:G + n[2], G:G + n[1], G:G + n[0]], g.to_dense(b))


def test_ghost_refresh_is_periodic_image():
    nb = (3, 2, 4)
    g = O.Grid(nb)
    a = make(g, 1)
    g.refresh(a)
    S0, S1, S2 = g.padded
    for k in range(S2):
        for j in range(S1):
            for i in range(S0):
                b = i + S0 * (j + S1 * k)
                si = (i - 1) % nb[0] + 1
                sj = (j - 1) % nb[1] + 1
                sk = (k - 1) % nb[2] + 1
                np.testing.assert_array_equal(a[b], a[si + S0 * (sj + S1 * sk)])


def test_fma_and_reference_order_agree_to_rounding():
    g = O.Grid((4, 4, 4))
    a = make(g, 2)
    g.refresh(a)
    for st in STENCILS:
        coeff = S.default_coeff(st, 3)
        b1, b2 = g.storage(), g.storage()
        g.sweep(st, coeff, a, b1, mode=O.MODE_FMA)
        g.sweep(st, coeff, a, b2, mode=O.MODE_REF)
        d = np.abs(g.to_dense(b1) - g.to_dense(b2)).max()
        assert d < 1e-15, (st, d)


def test_planted_mistakes_are_caught():
    """Checks with teeth: a swapped tap pair, a dropped tap and a sign flip
    in the oracle's coefficients each break the brute-force agreement."""
    g = O.Grid((2, 2, 2))
    a = make(g, 4)
    g.refresh(a)
    coeff = np.random.default_rng(1).uniform(0, 1, 7)
    exp = dense_ref(g.to_dense(a), O.STAR7, coeff)
    for bad in ("swap", "drop", "sign"):
        c = coeff.copy()
        if bad == "swap":
            c[1], c[2] = c[2], c[1]
        elif bad == "drop":
            c[5] = 0.0
        else:
            c[3] = -c[3]
        b = g.storage()
        g.sweep(O.STAR7, c, a, b)
        assert np.abs(g.to_dense(b) - exp).max() > 1e-3, bad


def test_hash_matches_python():
    for seed, n in [(0, 0), (1, 5), (12345, 1 << 40), ((1 << 64) - 1, 7)]:
        assert O.lib().or_hash64(seed, n) == S.hash64(seed, n)
        assert O.lib().or_u01(seed, n) == S.u01(seed, n)


def test_fill_matches_dense_generator():
    nb = (2, 1, 3)
    g = O.Grid(nb)
    a = make(g, 77)
    n = [8 * v for v in nb]
    arr = np.empty(int(np.prod(n)))
    O.lib().or_fill_uniform_array(O.P(arr, O.dblp), O.L3(n), 77, -1.0, 1.0)
    np.testing.assert_array_equal(g.to_dense(a).reshape(-1), arr)


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 1e-40, 65504.0, np.inf, -np.inf],
                    np.float32)
    got = np.array([O.lib().or_f32_to_bf16(float(v)) for v in vals], np.uint16)
    np.testing.assert_array_equal(got, O.f32_to_bf16(vals))
    # ties go to even: 1 + 2^-8 is halfway between bf16 1.0 and 1.0078125
    assert O.f32_to_bf16(np.array([1.00390625], np.float32))[0] == 0x3F80
    back = O.bf16_to_f32(got)
    assert np.all(np.abs(back[:4] - vals[:4]) <= np.abs(vals[:4]) * 2 ** -8)


@pytest.mark.parametrize("name,gi", [("7pt", 8), ("13pt", 4), ("25pt", 2), ("125pt", 4),
                                     ("cond", 8)])
def test_oracle_deep_ghost_loop_interior_exact(name, gi):
    """The communication-avoiding loop (weak/main.cpp:26-36,203-211) keeps the
    interior exact while radius * ghost_iters <= 8: stale ghost values move
    inwards one radius per sweep and never reach the interior before the
    next refresh.  One sweep more per group does reach it (checked for the
    linear stencils; cond's select can hide a stale value)."""
    from bricklib_amd import stencils as S
    nb = (3, 2, 4)
    st = S.Stencil.make(name, seed=23)
    g = O.Grid(nb)
    n = 2 * gi + 3
    outs = []
    for k in (1, gi):
        a, b = g.storage(), g.storage()
        g.fill_uniform(a, 5)
        outs.append(g.deep_sweeps(st.kind, st.coeff, a, b, n, k, theta=st.theta,
                                  cond_mode=st.cond_mode)[g.interior])
    np.testing.assert_array_equal(outs[1], outs[0])
    if name == "cond":
        return
    a, b = g.storage(), g.storage()
    g.fill_uniform(a, 5)
    over = g.deep_sweeps(st.kind, st.coeff, a, b, n, gi + 1, theta=st.theta,
                         cond_mode=st.cond_mode)[g.interior]
    assert not np.array_equal(over, outs[0])
