"""ORACLE / TEST INFRASTRUCTURE ONLY: never imported by the product path.

Evaluate SymPy expressions at the reference's md5-seeded arbitrary-precision
inputs, so this repository's symbolic transcription can be compared with the
reference's trusted values.

Input rule (restated from ``nrpy/validate_expressions/validate_expressions.py:368-400``):
every free symbol ``s`` gets ``random.seed(int(md5(str(s)).hexdigest(), 16));
mpf(random.random())``; evaluation happens at 30 significant digits
(``validate_expressions.py:22``) and agreement is judged with relative
tolerance ``10**(-0.8*30)`` (``validate_expressions.py:699-701``).
"""

from __future__ import annotations

import hashlib
import random
from typing import Dict

import mpmath
import sympy as sp

PRECISION = 30


def seeded_value(name: str) -> mpmath.mpf:
    """The md5-seeded input value the reference assigns to a symbol name."""
    if name in ("PI", "M_PI"):
        return mpmath.mpf(mpmath.pi)
    if name in ("SQRT1_2", "M_SQRT1_2"):
        return mpmath.mpf(1) / mpmath.sqrt(2)
    random.seed(int(hashlib.md5(name.encode()).hexdigest(), 16))
    return mpmath.mpf(random.random())


def evaluate(expr: sp.Expr, dps: int = PRECISION):
    """Evaluate an expression at md5-seeded inputs with ``dps`` digits."""
    expr = sp.sympify(expr)
    if expr == 0:
        return mpmath.mpf(0)
    old = mpmath.mp.dps
    mpmath.mp.dps = dps
    try:
        replaced, reduced = sp.cse(expr, order="none")
        env: Dict[sp.Symbol, object] = {}
        syms = set(reduced[0].free_symbols)
        for _, sub in replaced:
            syms |= sub.free_symbols
        for s in syms:
            env[s] = seeded_value(str(s))
        for lhs, rhs in replaced:
            env[lhs] = _num(rhs, env, dps)
        return _num(reduced[0], env, dps)
    finally:
        mpmath.mp.dps = old


_FUNCS = {
    sp.exp: mpmath.exp,
    sp.sin: mpmath.sin,
    sp.cos: mpmath.cos,
    sp.sqrt: mpmath.sqrt,
    sp.Abs: abs,
    sp.log: mpmath.log,
    sp.sinh: mpmath.sinh,
    sp.cosh: mpmath.cosh,
}


def _num(e: sp.Expr, env, dps: int):
    """Evaluate a SymPy tree with mpmath arithmetic (no floating round trips)."""
    if e in env:
        return env[e]
    if e.is_Integer:
        return mpmath.mpf(int(e))
    if e.is_Rational:
        return mpmath.mpf(int(e.p)) / int(e.q)
    if e.is_Number:
        return mpmath.mpf(sp.N(e, dps + 10))
    if e is sp.pi:
        return +mpmath.pi
    if e.is_Add:
        acc = mpmath.mpf(0)
        for a in e.args:
            acc += _num(a, env, dps)
        return acc
    if e.is_Mul:
        acc = mpmath.mpf(1)
        for a in e.args:
            acc *= _num(a, env, dps)
        return acc
    if e.is_Pow:
        base = _num(e.args[0], env, dps)
        ex = e.args[1]
        if ex.is_Integer:
            return base ** int(ex)
        if ex.is_Rational:
            # real principal root, matching C's cbrt/sqrt for positive bases
            return mpmath.power(base, mpmath.mpf(int(ex.p)) / int(ex.q))
        return mpmath.power(base, _num(ex, env, dps))
    for fn, mfn in _FUNCS.items():
        if isinstance(e, fn):
            return mfn(_num(e.args[0], env, dps))
    raise TypeError(f"cannot evaluate {type(e)}: {e}")


def agrees(value, trusted, dps: int = PRECISION) -> bool:
    """Relative agreement test used by the reference's validator."""
    mpmath.mp.dps = dps
    tol = mpmath.mpf(10) ** (-0.8 * dps)
    if isinstance(trusted, (list, tuple)):
        trusted = mpmath.mpc(trusted[0], trusted[1])
    else:
        trusted = mpmath.mpf(trusted)
    diff = abs(value - trusted)
    if trusted == 0:
        return diff < tol
    return diff <= abs(trusted) * tol
