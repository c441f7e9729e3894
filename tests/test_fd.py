"""Finite-difference stencils: the reference's doctest values, exactness on
polynomials, convergence order, and the Kreiss-Oliger sign convention."""

import numpy as np
import pytest
import sympy as sp

from nrpy_amd.codegen import fd


def test_reference_doctest_values():
    """nrpy/finite_difference.py doctests (:105-110, :1000-1011) at fd_order 2."""
    assert fd.weights_1d("dD", 2) == [(-1, sp.Rational(-1, 2)), (1, sp.Rational(1, 2))]
    assert fd.weights_1d("dupD", 2) == [(0, sp.Rational(-3, 2)), (1, 2), (2, sp.Rational(-1, 2))]
    assert fd.weights_1d("ddnD", 2) == [(-2, sp.Rational(1, 2)), (-1, -2), (0, sp.Rational(3, 2))]
    assert fd.weights_1d("dKOD", 2) == [
        (-2, sp.Rational(-1, 16)), (-1, sp.Rational(1, 4)), (0, sp.Rational(-3, 8)),
        (1, sp.Rational(1, 4)), (2, sp.Rational(-1, 16)),
    ]
    # mixed second derivative = product of centred first-derivative stencils
    pts = fd.parse_deriv("FDPROTO_dDD01").points(2)
    assert pts == {(-1, -1, 0): sp.Rational(1, 4), (1, -1, 0): sp.Rational(-1, 4),
                   (-1, 1, 0): sp.Rational(-1, 4), (1, 1, 0): sp.Rational(1, 4)}


@pytest.mark.parametrize("order", [2, 4, 6, 8])
@pytest.mark.parametrize("op,deriv", [("dD", 1), ("dDD", 2), ("dupD", 1), ("ddnD", 1)])
def test_exact_on_polynomials(order, op, deriv):
    """A stencil of accuracy N differentiates x^k exactly for k <= N (at x = 0)."""
    w = fd.weights_1d(op, order)
    for k in range(order + 1):
        val = sum(c * sp.Integer(o) ** k for o, c in w)
        exact = sp.factorial(k) if k == deriv else 0
        assert val == exact, (op, order, k)


@pytest.mark.parametrize("order", [2, 4, 6])
def test_kreiss_oliger(order):
    """KO annihilates polynomials below degree order+2 and damps the grid-scale mode
    with coefficient -1 (so eps * KO is dissipative)."""
    w = fd.weights_1d("dKOD", order)
    for k in range(order + 2):
        assert sum(c * sp.Integer(o) ** k for o, c in w) == 0
    nyquist = sum(c * sp.Integer(-1) ** o for o, c in w)
    assert nyquist == -1


@pytest.mark.parametrize("order", [4, 6])
def test_convergence_order(order):
    """Error of d/dx sin(x) falls as h^order."""
    errs = []
    for h in (0.1, 0.05):
        x0 = 0.3
        approx = sum(float(c) * np.sin(x0 + o * h) for o, c in fd.weights_1d("dD", order)) / h
        errs.append(abs(approx - np.cos(x0)))
    rate = np.log2(errs[0] / errs[1])
    assert abs(rate - order) < 0.3


def test_parse_derivative_names():
    d = fd.parse_deriv("hDD_dDD0112")
    assert (d.gf, d.op, d.dirs) == ("hDD01", "dDD", (1, 2))
    d = fd.parse_deriv("vetU_dupD21")
    assert (d.gf, d.op, d.dirs) == ("vetU2", "dupD", (1,))
    d = fd.parse_deriv("alpha_dKOD0")
    assert (d.gf, d.op, d.dirs) == ("alpha", "dKOD", (0,))
    with pytest.raises(ValueError):
        fd.parse_deriv("hDD_dD01")  # missing index


def test_upwind_union_stencil_matches_one_sided():
    """The kernels' per-direction weight vector (selected by sign of beta) equals the
    one-sided dupD / ddnD stencils on the union of their points."""
    for order in (4, 6):
        R = fd.ghost_zones(order)
        up = dict(fd.weights_1d("dupD", order))
        dn = dict(fd.weights_1d("ddnD", order))
        assert set(up) | set(dn) <= set(range(-R, R + 1))
        f = np.random.default_rng(0).normal(size=2 * R + 1)
        for U, ref in ((1.0, up), (0.0, dn)):
            sel = [float(up.get(k, 0)) if U > 0 else float(dn.get(k, 0)) for k in range(-R, R + 1)]
            a = sum(s * f[k + R] for k, s in zip(range(-R, R + 1), sel))
            b = sum(float(c) * f[k + R] for k, c in ref.items())
            assert abs(a - b) < 1e-14
