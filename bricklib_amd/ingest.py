"""Stencil descriptions in the codegen DSL, read as text.

A reference call site ``brick("X.py", VSVEC, (8,8,8), (8,8), b)``
(include/vecscatter.h:62-90) names a stencil description file written with
the codegen's expression classes -- ``Index``, ``Grid``, ``ConstRef``,
``Func``, ``If`` (codegen/st/expr.py, codegen/st/grid.py, codegen/st/func.py;
examples stencils/7pt.py, cond.py, mpi13pt.py, mpi25pt.py, mpi125pt.py).
This module parses such a file with Python's ``ast`` module -- the file is
never executed -- and evaluates the small DSL symbolically into a list of
taps (grid offset, coefficient symbol, per-tap max(x, 0)) plus the output
select, then maps it onto a kernel of this library:

  * taps {0} + {+-r e_a : r = 1..R} with R in {1, 2, 4}: the 7/13/25-point
    star (BK_STAR7/13/25);
  * the same 7 taps with max(x, 0) on each and the output
    ``If(calc > 0, calc, -calc)``: cond in its reference form (BK_COND7);
  * all 125 offsets of [-2, 2]^3: the 125-point cube (BK_CUBE125).

Anything else (other radii, missing or repeated taps, grids of other
dimensions such as stencils/mpi9pt.py's 4-D grid, other functions) is
rejected with ValueError.  The result gives the kernel-order coefficient
vector from the values of the file's ConstRef symbols: ``coeff[n]``
references read a coefficient array as the reference's drivers do
(stencils/3axis.cpp:47-66), named constants (MPI_B0, ...) come from a dict,
numeric ConstRefs ("0.2") are their value.  The kernels accumulate in their
own fixed order (centre, then +-r per axis); the file's textual order is
what the oracle's reference mode follows.

Command line: ``python -m bricklib_amd.ingest FILE.py`` prints the mapping.
"""
from __future__ import annotations

import ast
import json
import re
import sys
from dataclasses import dataclass, field

import numpy as np

from ._lib import BK_COND7, BK_CUBE125, BK_STAR7, BK_STAR13, BK_STAR25
from .stencils import Stencil, cube_offsets, star_offsets

_NUM = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?$")
_COEFF = re.compile(r"^(\w+)\[(\d+)\]$")


# -- symbolic values of the DSL ------------------------------------------------
@dataclass(frozen=True)
class _Index:
    axis: int
    shift: int = 0


@dataclass(frozen=True)
class _Grid:
    name: str
    ndim: int


@dataclass(frozen=True)
class _Func:
    name: str
    nargs: int


@dataclass(frozen=True)
class _Const:
    """A ConstRef: a named symbol or a numeric literal."""
    text: str

    @property
    def value(self):
        return float(self.text) if _NUM.match(self.text) else None


@dataclass(frozen=True)
class _Access:
    grid: _Grid
    offset: tuple  # per axis, axis 0 first (i, j, k, ...)


@dataclass
class _Lin:
    """A linear combination of taps: [(coefficient, access, relu)]."""
    terms: list = field(default_factory=list)


@dataclass(frozen=True)
class _Relu:
    access: _Access


@dataclass
class _Select:
    """If(calc > 0, calc, -calc) over a linear combination."""
    lin: _Lin


@dataclass
class _Cmp:
    lhs: object
    op: str
    rhs: object


@dataclass
class _Neg:
    value: object


@dataclass
class Tap:
    offset: tuple  # (di, dj, dk)
    coeff: str     # ConstRef text
    relu: bool = False


@dataclass
class StencilSpec:
    """A parsed stencil description mapped onto a kernel."""
    path: str
    input: str
    output: str
    taps: list          # Tap, in the file's textual order
    select_abs: bool    # output If(calc > 0, calc, -calc)
    kind: int           # BK_STAR7 ... BK_COND7
    order: list         # coefficient symbol of each kernel tap, kernel order

    @property
    def symbols(self) -> list:
        """Distinct coefficient symbols in first-use order (numeric literals
        excluded)."""
        seen = []
        for t in self.taps:
            if _Const(t.coeff).value is None and t.coeff not in seen:
                seen.append(t.coeff)
        return seen

    def coefficients(self, values=None) -> np.ndarray:
        """Kernel-order coefficients.  values: a sequence for ``name[n]``
        symbols (the coeff array of the reference drivers), a dict
        {symbol: value} for named constants, or a dict holding both (an
        array under the array's name)."""
        out = np.empty(len(self.order))
        for q, sym in enumerate(self.order):
            out[q] = _resolve(sym, values)
        return out

    def stencil(self, values=None) -> Stencil:
        """The Stencil (kind + coefficients) the sweep entry points take;
        cond in its reference form."""
        return Stencil(kind=self.kind, coeff=self.coefficients(values), theta=0.0,
                       cond_mode=0)


def _resolve(sym: str, values) -> float:
    lit = _Const(sym).value
    if lit is not None:
        return lit
    m = _COEFF.match(sym)
    if isinstance(values, dict):
        if sym in values:
            return float(values[sym])
        if m and m.group(1) in values:
            return float(values[m.group(1)][int(m.group(2))])
    elif values is not None and m:
        return float(values[int(m.group(2))])
    raise KeyError(f"no value for coefficient symbol {sym!r}")


# -- the evaluator ----------------------------------------------------------------
class _Eval:
    ALLOWED_IMPORTS = {"st.expr", "st.grid", "st.func"}

    def __init__(self, path: str):
        self.path = path
        self.env: dict = {}
        self.assigned = None  # (output access, value)

    def fail(self, node, msg):
        line = getattr(node, "lineno", "?")
        raise ValueError(f"{self.path}:{line}: {msg}")

    def run(self, tree: ast.Module):
        for st in tree.body:
            if isinstance(st, ast.ImportFrom):
                if st.module not in self.ALLOWED_IMPORTS:
                    self.fail(st, f"import from {st.module!r} is not part of the stencil DSL")
                for a in st.names:
                    self.env[a.asname or a.name] = ("ctor", a.name)
            elif isinstance(st, ast.Assign):
                val = self.expr(st.value)
                for t in st.targets:
                    if not isinstance(t, ast.Name):
                        self.fail(st, "only plain names can be assigned")
                    self.env[t.id] = val
            elif isinstance(st, ast.Expr):
                self.statement_call(st.value)
            else:
                self.fail(st, f"unsupported statement {type(st).__name__}")

    def statement_call(self, node):
        # output(i, j, k).assign(expr)
        if (isinstance(node, ast.Call) and isinstance(node.func, ast.Attribute)
                and node.func.attr == "assign" and len(node.args) == 1):
            target = self.expr(node.func.value)
            if not isinstance(target, _Access):
                self.fail(node, "assign() on something that is not a grid access")
            if self.assigned is not None:
                self.fail(node, "more than one output assignment")
            self.assigned = (target, self.expr(node.args[0]))
        else:
            self.fail(node, "the only statement call allowed is grid(...).assign(expr)")

    def expr(self, n):
        if isinstance(n, ast.Constant):
            return n.value
        if isinstance(n, ast.Name):
            if n.id not in self.env:
                self.fail(n, f"unknown name {n.id!r}")
            return self.env[n.id]
        if isinstance(n, ast.List):
            return [self.expr(e) for e in n.elts]
        if isinstance(n, ast.Subscript):
            base = self.expr(n.value)
            idx = self.expr(n.slice)
            if not isinstance(base, list) or not isinstance(idx, int):
                self.fail(n, "only list[int] subscripts are supported")
            return base[idx]
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, ast.USub):
            v = self.expr(n.operand)
            if isinstance(v, (int, float)):
                return -v
            return _Neg(v)
        if isinstance(n, ast.BinOp):
            return self.binop(n, self.expr(n.left), self.expr(n.right))
        if isinstance(n, ast.Compare):
            if len(n.ops) != 1 or not isinstance(n.ops[0], ast.Gt):
                self.fail(n, "only `a > b` comparisons are supported")
            return _Cmp(self.expr(n.left), ">", self.expr(n.comparators[0]))
        if isinstance(n, ast.Call):
            return self.call(n)
        self.fail(n, f"unsupported expression {type(n).__name__}")

    def call(self, n):
        f = self.expr(n.func)
        args = [self.expr(a) for a in n.args]
        if isinstance(f, tuple) and f[0] == "ctor":
            name = f[1]
            if name == "Index" and len(args) == 1 and isinstance(args[0], int):
                return _Index(args[0])
            if name == "Grid" and len(args) == 2 and isinstance(args[0], str):
                return _Grid(args[0], int(args[1]))
            if name == "ConstRef" and len(args) == 1 and isinstance(args[0], str):
                return _Const(args[0])
            if name == "Func" and len(args) == 2 and isinstance(args[0], str):
                return _Func(args[0], int(args[1]))
            if name == "If" and len(args) == 3:
                return self.if_expr(n, *args)
            self.fail(n, f"unsupported constructor call {name}{tuple(args)!r}")
        if isinstance(f, _Grid):
            if len(args) != f.ndim or not all(isinstance(a, _Index) for a in args):
                self.fail(n, f"grid {f.name!r} needs {f.ndim} index arguments")
            axes = [a.axis for a in args]
            if axes != list(range(f.ndim)):
                self.fail(n, "grid accesses must use the indices in axis order (i, j, k, ...)")
            return _Access(f, tuple(a.shift for a in args))
        if isinstance(f, _Func):
            if f.name == "max" and len(args) == 2 and isinstance(args[0], _Access) \
                    and isinstance(args[1], _Const) and args[1].value == 0.0:
                return _Relu(args[0])
            self.fail(n, f"only max(grid access, 0) is supported, not {f.name}")
        self.fail(n, "call of something that is not callable in the DSL")

    def if_expr(self, n, cond, a, b):
        # If(calc > 0, calc, -calc): |calc| with the sign rule of cond.py
        if (isinstance(cond, _Cmp) and isinstance(cond.lhs, _Lin) and cond.lhs is a
                and isinstance(cond.rhs, (int, float)) and cond.rhs == 0
                and isinstance(b, _Neg) and b.value is a):
            return _Select(a)
        self.fail(n, "only If(calc > 0, calc, -calc) is supported")

    def binop(self, n, a, b):
        if isinstance(a, _Index) and isinstance(b, int):
            if isinstance(n.op, ast.Add):
                return _Index(a.axis, a.shift + b)
            if isinstance(n.op, ast.Sub):
                return _Index(a.axis, a.shift - b)
        if isinstance(n.op, ast.Mult):
            if isinstance(a, _Const) and isinstance(b, (_Access, _Relu)):
                return self.term(a, b)
            if isinstance(b, _Const) and isinstance(a, (_Access, _Relu)):
                return self.term(b, a)
        if isinstance(n.op, ast.Add) and isinstance(a, _Lin) and isinstance(b, _Lin):
            return _Lin(a.terms + b.terms)
        self.fail(n, "expressions must be sums of ConstRef * grid access")

    @staticmethod
    def term(c: _Const, x):
        if isinstance(x, _Relu):
            return _Lin([(c.text, x.access, True)])
        return _Lin([(c.text, x, False)])


def _classify(path, taps, select_abs):
    offs = [t.offset for t in taps]
    if len(set(offs)) != len(offs):
        raise ValueError(f"{path}: a grid offset appears in more than one term")
    by_off = {t.offset: t for t in taps}
    relus = {t.relu for t in taps}
    if len(relus) != 1:
        raise ValueError(f"{path}: max(x, 0) must apply to every tap or to none")
    relu = relus.pop()
    candidates = [(BK_STAR7, star_offsets(1)), (BK_STAR13, star_offsets(2)),
                  (BK_STAR25, star_offsets(4)), (BK_CUBE125, cube_offsets(2))]
    for kind, ko in candidates:
        if set(ko) == set(offs):
            if relu or select_abs:
                if kind != BK_STAR7 or not (relu and select_abs):
                    raise ValueError(f"{path}: max(x, 0) taps with the If(calc > 0, calc, "
                                     f"-calc) output exist only for the 7-point star (cond)")
                kind = BK_COND7
            return kind, [by_off[o].coeff for o in ko]
    raise ValueError(f"{path}: taps {sorted(offs)} match no kernel (7/13/25-point star, "
                     f"125-point cube, cond)")


def parse_text(text: str, path: str = "<text>") -> StencilSpec:
    tree = ast.parse(text, filename=path)
    ev = _Eval(path)
    ev.run(tree)
    if ev.assigned is None:
        raise ValueError(f"{path}: no output(...).assign(...) statement")
    out_access, value = ev.assigned
    if out_access.grid.ndim != 3:
        raise ValueError(f"{path}: {out_access.grid.ndim}-D grids are not supported (3-D only)")
    if any(out_access.offset):
        raise ValueError(f"{path}: the output must be written at (i, j, k)")
    select_abs = isinstance(value, _Select)
    lin = value.lin if select_abs else value
    if not isinstance(lin, _Lin):
        raise ValueError(f"{path}: the assigned expression is not a sum of taps")
    grids = {acc.grid.name for _, acc, _ in lin.terms}
    if len(grids) != 1:
        raise ValueError(f"{path}: taps read {len(grids)} grids; one input grid is supported")
    if out_access.grid.name in grids:
        raise ValueError(f"{path}: the output grid is also read")
    taps = [Tap(offset=acc.offset, coeff=c, relu=r) for c, acc, r in lin.terms]
    kind, order = _classify(path, taps, select_abs)
    return StencilSpec(path=path, input=grids.pop(), output=out_access.grid.name, taps=taps,
                       select_abs=select_abs, kind=kind, order=order)


def parse_file(path: str) -> StencilSpec:
    with open(path) as f:
        return parse_text(f.read(), path)


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("usage: python -m bricklib_amd.ingest FILE.py [FILE.py ...]", file=sys.stderr)
        return 2
    rc = 0
    names = {BK_STAR7: "BK_STAR7", BK_STAR13: "BK_STAR13", BK_STAR25: "BK_STAR25",
             BK_CUBE125: "BK_CUBE125", BK_COND7: "BK_COND7"}
    for p in argv:
        try:
            sp = parse_file(p)
        except (ValueError, SyntaxError) as e:
            print(json.dumps({"file": p, "error": str(e)}))
            rc = 1
            continue
        print(json.dumps({"file": p, "kernel": names[sp.kind], "input": sp.input,
                          "output": sp.output, "symbols": sp.symbols,
                          "coefficient_order": sp.order}))
    return rc


if __name__ == "__main__":
    sys.exit(main())
