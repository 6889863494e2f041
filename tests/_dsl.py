"""Stencil descriptions written in the codegen DSL (the expression classes
of codegen/st/expr.py, grid.py, func.py, as the reference's stencils/*.py
files use them), composed here from (offset, coefficient) taps for the
ingestion tests."""


HEAD = """from st.expr import Index, ConstRef, If
from st.grid import Grid
from st.func import Func
i = Index(0)
j = Index(1)
k = Index(2)
src = Grid("u", 3)
dst = Grid("v", 3)
"""


def acc(d):
    di, dj, dk = d
    f = lambda n, v: n if v == 0 else (f"{n} + {v}" if v > 0 else f"{n} - {-v}")  # noqa: E731
    return f"src({f('i', di)}, {f('j', dj)}, {f('k', dk)})"


def text_of(taps, relu=False, select=False):
    """DSL text for [(offset, const_name_or_literal)] in the given order."""
    lines = [HEAD, "maxf = Func(\"max\", 2)", "zero = ConstRef(\"0.0\")"]
    names = {}
    for _, c in taps:
        if c not in names:
            names[c] = f"q{len(names)}"
            lines.append(f"{names[c]} = ConstRef(\"{c}\")")
    terms = []
    for d, c in taps:
        x = f"maxf({acc(d)}, zero)" if relu else acc(d)
        terms.append(f"{names[c]} * {x}")
    lines.append("calc = " + " + \\\n    ".join(terms))
    if select:
        lines.append("calc = If(calc > 0, calc, -calc)")
    lines.append("dst(i, j, k).assign(calc)")
    lines.append("STENCIL = [dst]")
    return "\n".join(lines) + "\n"



def star(R):
    """The star offsets (di, dj, dk) of radius R in kernel coefficient order."""
    offs = [(0, 0, 0)]
    for r in range(1, R + 1):
        offs += [(r, 0, 0), (-r, 0, 0), (0, r, 0), (0, -r, 0), (0, 0, r), (0, 0, -r)]
    return offs
