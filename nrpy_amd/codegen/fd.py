"""Finite-difference stencils and derivative-symbol parsing.

Coefficients are exact rationals from the Vandermonde system used by the
reference (``nrpy/finite_difference.py:93-131`` and ``:202-330``):

* ``dD``   centred first derivative, width fd_order+1
* ``dDD``  centred second derivative (unmixed), width fd_order+1; mixed
  second derivatives are products of two centred first-derivative stencils
* ``dupD`` / ``ddnD`` first derivative on a stencil shifted by +1 / -1
* ``dKOD`` Kreiss-Oliger operator: the (fd_order+2)-th difference on
  fd_order+3 points, times ``(-1)**((w+1)/2) / 2**(fd_order+2)`` and a
  single factor ``1/dx`` (``finite_difference.py:225-226, 288-291, 1062-1068``)

Derivative symbols are parsed the reference's way
(``finite_difference.py:473-573``): ``hDD_dDD0112`` is operator ``dDD`` of
gridfunction ``hDD01`` along directions (1, 2).
"""

from __future__ import annotations

import functools
import re
from dataclasses import dataclass
from typing import Dict, List, Tuple

import sympy as sp


@functools.lru_cache(maxsize=None)
def stencil_weights(width: int, shift: int, deriv: int) -> Tuple[Tuple[int, sp.Rational], ...]:
    """Weights of the ``deriv``-th derivative on ``width`` points shifted by ``shift``.

    Returns ((offset, weight), ...) for nonzero weights, for unit spacing.
    """
    half = sp.Rational(width - 1, 2)
    offs = [j - half + shift for j in range(width)]
    M = sp.Matrix(width, width, lambda i, j: sp.Integer(1) if i == 0 else offs[j] ** i)
    rhs = sp.zeros(width, 1)
    rhs[deriv] = sp.factorial(deriv)
    c = M.LUsolve(rhs)
    return tuple((int(offs[j]), sp.nsimplify(c[j])) for j in range(width) if c[j] != 0)


def weights_1d(op: str, fd_order: int) -> List[Tuple[int, sp.Rational]]:
    """1D weights for ``dD``, ``dDD`` (unmixed), ``dupD``, ``ddnD``, ``dKOD``."""
    if op == "dD":
        return list(stencil_weights(fd_order + 1, 0, 1))
    if op == "dDD":
        return list(stencil_weights(fd_order + 1, 0, 2))
    if op == "dupD":
        return list(stencil_weights(fd_order + 1, 1, 1))
    if op == "ddnD":
        return list(stencil_weights(fd_order + 1, -1, 1))
    if op == "dKOD":
        w = fd_order + 3
        k = fd_order + 2
        sign = sp.Integer(-1) ** sp.Rational(w + 1, 2)
        return [(o, c * sign / 2**k) for o, c in stencil_weights(w, 0, k)]
    raise ValueError(op)


def ghost_zones(fd_order: int, upwind_or_KO: bool = True) -> int:
    """NGHOSTS = fd_order/2, plus one with upwinding or KO (BHaH_defines_h.py:278-292)."""
    return fd_order // 2 + (1 if upwind_or_KO else 0)


@dataclass(frozen=True)
class Deriv:
    """One derivative symbol: operator applied to a gridfunction along dirs."""

    symbol: str
    gf: str
    op: str  # dD, dDD, dupD, ddnD, dKOD
    dirs: Tuple[int, ...]

    def points(self, fd_order: int) -> Dict[Tuple[int, int, int], sp.Rational]:
        """3D stencil: {(di0, di1, di2): weight} for unit spacing."""
        pts: Dict[Tuple[int, int, int], sp.Rational] = {}
        if self.op == "dDD" and self.dirs[0] != self.dirs[1]:
            w = weights_1d("dD", fd_order)
            for oa, ca in w:
                for ob, cb in w:
                    o = [0, 0, 0]
                    o[self.dirs[0]] += oa
                    o[self.dirs[1]] += ob
                    pts[tuple(o)] = pts.get(tuple(o), 0) + ca * cb
            return pts
        for off, c in weights_1d(self.op, fd_order):
            o = [0, 0, 0]
            o[self.dirs[0]] = off
            pts[tuple(o)] = c
        return pts


_DERIV_RE = re.compile(r"^(?P<base>[A-Za-z]+?)(?P<ud>[UD]*)_(?P<op>dKOD|dupD|ddnD|dDD|dD)(?P<digits>\d+)$")


def parse_deriv(name: str) -> Deriv:
    """Parse a derivative symbol name into gridfunction, operator, directions."""
    m = _DERIV_RE.match(name)
    if not m:
        raise ValueError(f"not a derivative symbol: {name}")
    rank = len(m.group("ud"))
    order = 2 if m.group("op") == "dDD" else 1
    digits = m.group("digits")
    if len(digits) != rank + order:
        raise ValueError(f"index count mismatch in {name}")
    gf = m.group("base") + m.group("ud") + digits[:rank]
    dirs = tuple(int(c) for c in digits[rank:])
    return Deriv(name, gf, m.group("op"), dirs)


def is_deriv_symbol(name: str) -> bool:
    return bool(_DERIV_RE.match(name))
