"""Symbolic tensor declarations with the reference's component-naming rules.

Component symbols are named ``<base><i><j>...`` exactly as the reference's
indexed-expression module names them (``nrpy/indexedexp.py:210-369``):
symmetric index pairs carry the smaller index first (``hDD10`` -> ``hDD01``,
``hDD_dDD1002`` -> ``hDD_dDD0102``).  The names matter beyond cosmetics: the
reference's trusted-value dictionaries seed each free symbol with
``md5(name)``, so the same names are needed to evaluate this transcription at
those inputs.

Derivative symbols (``..._dD..``, ``..._dDD..``, ``..._dupD..``,
``..._dKOD..``) whose derivative direction lies along a symmetry axis are
replaced by zero, following ``nrpy/indexedexp.py:598-700``.
"""

from __future__ import annotations

import itertools
from typing import List, Sequence

import sympy as sp

# Symmetry axes (string of digits, e.g. "2" for axisymmetry about phi in
# spherical-like coordinates).  Module-level state, set by the caller before
# building expressions; mirrors the reference's ``symmetry_axes`` parameter.
_SYMMETRY_AXES = ""


def set_symmetry_axes(axes: str) -> None:
    """Set the coordinate directions along which derivatives vanish."""
    global _SYMMETRY_AXES
    _SYMMETRY_AXES = axes


def get_symmetry_axes() -> str:
    """Return the current symmetry-axes string."""
    return _SYMMETRY_AXES


def _deriv_order(name: str) -> int:
    """Number of derivative indices at the end of a derivative symbol name."""
    pos = name.rfind("_d")
    if pos < 0:
        return 0
    tail = name[pos + 2 :]
    if tail.startswith("DDD"):
        raise ValueError(f"derivative order > 2 not supported: {name}")
    return 2 if tail.startswith("DD") else 1


def _crosses_symmetry_axis(name: str) -> bool:
    order = _deriv_order(name)
    if order == 0 or not _SYMMETRY_AXES:
        return False
    return any(c in _SYMMETRY_AXES for c in name[len(name) - order :])


def _canonical(idx: Sequence[int], symmetry: str) -> List[int]:
    """Sort symmetric index pairs so the smaller index comes first."""
    out = list(idx)
    if symmetry in ("", "nosym"):
        return out
    for token in symmetry.split("_"):
        if not token.startswith("sym"):
            raise ValueError(f"unsupported symmetry token {token}")
        pos = [int(c) for c in token[3:]]
        vals = sorted(out[p] for p in pos)
        for p, v in zip(pos, vals):
            out[p] = v
    return out


def declare(base: str, rank: int, symmetry: str = "nosym", dim: int = 3):
    """Declare a rank-1..4 tensor of symbols as nested lists.

    Derivatives across symmetry axes become ``sp.Integer(0)``.
    """

    def build(prefix: List[int]):
        if len(prefix) == rank:
            idx = _canonical(prefix, symmetry)
            name = base + "".join(str(i) for i in idx)
            if _crosses_symmetry_axis(name):
                return sp.Integer(0)
            return sp.Symbol(name, real=True)
        return [build(prefix + [i]) for i in range(dim)]

    return build([])


def zeros(rank: int, dim: int = 3):
    """Nested lists of sympy zeros."""
    if rank == 0:
        return sp.Integer(0)
    return [zeros(rank - 1, dim) for _ in range(dim)]


def idx(rank: int, dim: int = 3):
    """Iterate over all index tuples of the given rank."""
    return itertools.product(range(dim), repeat=rank)


def symm_inverse3x3(m):
    """Inverse and determinant of a symmetric 3x3 matrix (nested lists)."""
    det = (
        m[0][0] * m[1][1] * m[2][2]
        + 2 * m[0][1] * m[1][2] * m[0][2]
        - m[0][0] * m[1][2] ** 2
        - m[1][1] * m[0][2] ** 2
        - m[2][2] * m[0][1] ** 2
    )
    inv = zeros(2)
    inv[0][0] = (m[1][1] * m[2][2] - m[1][2] ** 2) / det
    inv[0][1] = inv[1][0] = (m[0][2] * m[1][2] - m[0][1] * m[2][2]) / det
    inv[0][2] = inv[2][0] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det
    inv[1][1] = (m[0][0] * m[2][2] - m[0][2] ** 2) / det
    inv[1][2] = inv[2][1] = (m[0][1] * m[0][2] - m[0][0] * m[1][2]) / det
    inv[2][2] = (m[0][0] * m[1][1] - m[0][1] ** 2) / det
    return inv, det
