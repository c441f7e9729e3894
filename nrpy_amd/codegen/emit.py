"""SymPy -> C/HIP expression emitter with common-subexpression elimination.

Prints double-precision C: integer powers become products, rationals become
named compile-time constants ``Rational_p_q = p.0/q.0`` (correctly rounded,
the same values the reference's ``static const REAL Rational_p_q`` hold),
``x**(1/3)`` becomes ``cbrt``, ``Abs`` becomes ``fabs``.
"""

from __future__ import annotations

from typing import Dict, Iterable, List, Sequence, Set, Tuple

import sympy as sp
from sympy.printing.c import C99CodePrinter


def rational_name(r: sp.Rational) -> str:
    p, q = int(r.p), int(r.q)
    sign = "m" if p < 0 else ""
    return f"Rational_{sign}{abs(p)}_{q}"


class DoublePrinter(C99CodePrinter):
    """C99 printer for double arithmetic with named rational constants."""

    def __init__(self, rationals: Dict[str, str]):
        super().__init__()
        self._rationals = rationals

    def _print_Rational(self, expr):
        if expr.q == 1:
            return f"{int(expr.p)}.0"
        name = rational_name(expr)
        self._rationals[name] = f"{int(expr.p)}.0 / {int(expr.q)}.0"
        return name

    def _print_Integer(self, expr):
        return f"{int(expr)}.0"

    def _print_Float(self, expr):
        return repr(float(expr))

    def _print_Pow(self, expr):
        b, e = expr.args
        PREC = 60
        bs = self.parenthesize(b, PREC)
        if e.is_Integer:
            n = int(e)
            if n == 1:
                return bs
            if n == -1:
                return f"(1.0 / {bs})"
            if 2 <= n <= 12:
                return "(" + "*".join([bs] * n) + ")"
            if -12 <= n <= -2:
                return "(1.0 / (" + "*".join([bs] * (-n)) + "))"
            return f"pow({self._print(b)}, {n}.0)"
        if e == sp.Rational(1, 2):
            return f"sqrt({self._print(b)})"
        if e == sp.Rational(-1, 2):
            return f"(1.0 / sqrt({self._print(b)}))"
        if e == sp.Rational(1, 3):
            return f"cbrt({self._print(b)})"
        if e == sp.Rational(-1, 3):
            return f"(1.0 / cbrt({self._print(b)}))"
        return f"pow({self._print(b)}, {self._print(e)})"

    def _print_Abs(self, expr):
        return f"fabs({self._print(expr.args[0])})"

    def _print_Function(self, expr):
        if expr.func.__name__ == "nrpyAbs":
            return f"fabs({self._print(expr.args[0])})"
        return super()._print_Function(expr)


def cse_code(
    lhs: Sequence[str],
    exprs: Sequence[sp.Expr],
    tmp_prefix: str = "t",
    indent: str = "  ",
    rationals: Dict[str, str] | None = None,
) -> Tuple[str, Dict[str, str]]:
    """Emit ``const double t_k = ...;`` CSE temporaries then ``lhs[i] = ...;``.

    ``lhs`` entries are complete C lvalues (or ``const double name``).
    Returns (code, rationals-used).
    """
    if rationals is None:
        rationals = {}
    pr = DoublePrinter(rationals)
    syms = sp.numbered_symbols(tmp_prefix)
    replaced, reduced = sp.cse(list(exprs), symbols=syms, order="none")
    lines: List[str] = []
    for s, e in replaced:
        lines.append(f"{indent}const double {s} = {pr.doprint(e)};")
    for l, e in zip(lhs, reduced):
        lines.append(f"{indent}{l} = {pr.doprint(e)};")
    return "\n".join(lines), rationals


def rational_decls(rationals: Dict[str, str], qualifier: str = "const double", indent: str = "  ") -> str:
    return "\n".join(f"{indent}{qualifier} {k} = {v};" for k, v in sorted(rationals.items()))


def count_flops(exprs: Iterable[sp.Expr]) -> int:
    """Count +,-,*,/ in a CSE'd expression list (pow by small ints as muls)."""
    replaced, reduced = sp.cse(list(exprs), order="none")
    total = 0
    for e in [r for _, r in replaced] + list(reduced):
        total += _flops(e)
    return total


def _flops(e) -> int:
    if e.is_Atom:
        return 0
    n = sum(_flops(a) for a in e.args)
    if e.is_Add or e.is_Mul:
        return n + len(e.args) - 1
    if e.is_Pow:
        ex = e.args[1]
        if ex.is_Integer:
            k = abs(int(ex))
            return n + max(k - 1, 0) + (1 if int(ex) < 0 else 0)
        return n + 1
    return n + 1


def free_symbol_names(exprs: Iterable[sp.Expr]) -> Set[str]:
    out: Set[str] = set()
    for e in exprs:
        out |= {str(s) for s in sp.sympify(e).free_symbols}
    return out
