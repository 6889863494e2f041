"""Stencil descriptions in the codegen DSL, parsed as text (never executed),
mapped onto the kernels.  The mapping is checked against a brute force that
applies the file's own taps (offset, coefficient) with numpy.roll, so a
transposed axis, a flipped sign or a shuffled coefficient fails."""
import os

import numpy as np
import pytest

import oracle as O
from bricklib_amd import ingest
from bricklib_amd._lib import BK_COND7, BK_CUBE125, BK_STAR7, BK_STAR13, BK_STAR25

from _dsl import HEAD, star, text_of

REF = "/root/reference/stencils"


def brute(field, spec, values, relu=False, select=False):
    out = np.zeros_like(field)
    for t in spec.taps:
        di, dj, dk = t.offset
        x = np.roll(field, shift=(-dk, -dj, -di), axis=(0, 1, 2))
        if relu:
            x = np.where(x < 0, 0.0, x)
        out += ingest._resolve(t.coeff, values) * x
    if select:
        out = np.where(out > 0, out, -out)
    return out


def oracle_sweep(spec, values, nb=(3, 2, 4)):
    g = O.Grid(nb)
    a, b = g.storage(), g.storage()
    g.fill_uniform(a, 99)
    inp = g.to_dense(a)
    st = spec.stencil(values)
    g.refresh(a)
    g.sweep(st.kind, st.coeff, a, b, mode=O.MODE_FMA, cond_mode=0)
    return inp, g.to_dense(b)


@pytest.mark.parametrize("R,kind", [(1, BK_STAR7), (2, BK_STAR13), (4, BK_STAR25)])
def test_shuffled_stars_map_and_compute(R, kind):
    rng = np.random.default_rng(R)
    offs = star(R)
    perm = rng.permutation(len(offs))
    taps = [(offs[p], f"c{p}") for p in perm]
    spec = ingest.parse_text(text_of(taps))
    assert spec.kind == kind and spec.input == "u" and spec.output == "v"
    values = {f"c{p}": float(rng.uniform(-1, 1)) for p in range(len(offs))}
    inp, got = oracle_sweep(spec, values)
    want = brute(inp, spec, values)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * np.abs(want).max())


def test_cube_with_symmetry_classes_and_literals():
    offs = [(di, dj, dk) for dk in range(-2, 3) for dj in range(-2, 3) for di in range(-2, 3)]
    rng = np.random.default_rng(7)
    taps = []
    for p in rng.permutation(125):
        d = offs[p]
        cls = tuple(sorted(abs(x) for x in d))
        taps.append((d, "0.125" if cls == (0, 0, 0) else f"C{cls[0]}{cls[1]}{cls[2]}"))
    spec = ingest.parse_text(text_of(taps))
    assert spec.kind == BK_CUBE125 and len(spec.symbols) == 9
    values = {s: float(rng.uniform(-0.1, 0.1)) for s in spec.symbols}
    co = spec.coefficients(values)
    assert co[62] == 0.125  # the centre, (0,0,0), in (dk, dj, di) order
    inp, got = oracle_sweep(spec, values)
    want = brute(inp, spec, values)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * np.abs(want).max())


def test_cond_reference_form():
    taps = [(d, f"coeff[{q}]") for q, d in enumerate(star(1))][::-1]
    spec = ingest.parse_text(text_of(taps, relu=True, select=True))
    assert spec.kind == BK_COND7
    coeff = np.random.default_rng(3).uniform(-1, 1, 7)
    np.testing.assert_array_equal(spec.coefficients(coeff), coeff)
    inp, got = oracle_sweep(spec, coeff)
    want = brute(inp, spec, coeff, relu=True, select=True)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * np.abs(want).max())


@pytest.mark.parametrize("mutate,msg", [
    (lambda t: t + [((3, 0, 0), "x")], "match no kernel"),              # radius 3 tap
    (lambda t: t[:-1], "match no kernel"),                              # a tap missing
    (lambda t: t + [t[0]], "more than one term"),                       # repeated offset
])
def test_unsupported_shapes_rejected(mutate, msg):
    taps = [(d, f"c{q}") for q, d in enumerate(star(1))]
    with pytest.raises(ValueError, match=msg):
        ingest.parse_text(text_of(mutate(taps)))


@pytest.mark.parametrize("body,msg", [
    ("import os\n", "unsupported statement"),
    ("from os import path\n", "not part of the stencil DSL"),
    ("__import__('os').system('true')\n", "only statement call"),
    ("def f():\n    pass\n", "unsupported statement"),
    ("w = Grid('w', 2)\nw(i, j).assign(ConstRef('a') * w(i, j))\n", "2-D grids"),
    ("dst(i, j, k).assign(ConstRef('a') * src(j, i, k))\n", "axis order"),
    ("dst(i, j, k).assign(ConstRef('a') * src(i, j, k) * src(i, j, k))\n", "sums of"),
])
def test_non_dsl_text_rejected_without_running(body, msg):
    with pytest.raises(ValueError, match=msg):
        ingest.parse_text(HEAD + body)


def test_missing_values_raise():
    spec = ingest.parse_text(text_of([(d, f"c{q}") for q, d in enumerate(star(1))]))
    with pytest.raises(KeyError):
        spec.coefficients({"c0": 1.0})


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")
def test_reference_stencil_files():
    want = {"7pt.py": BK_STAR7, "mpi7pt.py": BK_STAR7, "cond.py": BK_COND7,
            "mpi13pt.py": BK_STAR13, "mpi25pt.py": BK_STAR25, "mpi125pt.py": BK_CUBE125}
    for name, kind in want.items():
        spec = ingest.parse_file(os.path.join(REF, name))
        assert spec.kind == kind, name
    # 7pt.py / cond.py read coeff[0..6] in the kernel's own tap order
    assert ingest.parse_file(os.path.join(REF, "7pt.py")).order == [f"coeff[{q}]"
                                                                    for q in range(7)]
    # mpi125pt.py: ten constants by symmetry class sorted(|di|, |dj|, |dk|)
    spec = ingest.parse_file(os.path.join(REF, "mpi125pt.py"))
    classes = [(0, 0, 0), (0, 0, 1), (0, 0, 2), (0, 1, 1), (0, 1, 2), (0, 2, 2), (1, 1, 1),
               (1, 1, 2), (1, 2, 2), (2, 2, 2)]
    offs = [(di, dj, dk) for dk in range(-2, 3) for dj in range(-2, 3) for di in range(-2, 3)]
    for d, sym in zip(offs, spec.order):
        assert sym == f"MPI_C{classes.index(tuple(sorted(abs(x) for x in d)))}", d
    with pytest.raises(ValueError, match="4-D"):
        ingest.parse_file(os.path.join(REF, "mpi9pt.py"))
