"""Reference metrics for Cartesian, Spherical and SinhSpherical coordinates.

Restates the spherical-like branch of the reference's ``ReferenceMetric``
(``nrpy/reference_metric.py:1195-1366``) in "rfm_precompute" form
(``nrpy/reference_metric.py:493-760``): every hatted quantity is first
written in terms of generic functions ``f0_of_xx0(xx0)`` and
``f1_of_xx1(xx1)`` and their derivatives, which are then replaced by plain
symbols ``f0_of_xx0``, ``f0_of_xx0__D0``, ``f0_of_xx0__DD00``,
``f0_of_xx0__DDD000``, ``f1_of_xx1``, ``f1_of_xx1__D1``, ``f1_of_xx1__DD11``.
Derivatives whose closed form does not depend on a coordinate (for example
``d f0/dxx0 = 1`` in Spherical coordinates) are substituted by their values.

On device, each of those symbols is a 1D array over the xx0 or xx1 grid
(the reference's ``rfm_struct``), filled at setup from the closed forms in
``RefMetric.precompute_defs``.

Spherical-like scale factors: ``(f0', f0, f0*f1)`` with ``f0 = r(xx0)`` and
``f1 = sin(xx1)`` (``nrpy/reference_metric.py:1338-1357``).
"""

from __future__ import annotations

from typing import Dict, List

import sympy as sp

from . import tensors as T

PI = sp.pi


class RefMetric:
    """Hatted quantities of one coordinate system, in precompute form."""

    def __init__(self, coord: str):
        self.coord = coord
        self.xx = [sp.Symbol(f"xx{i}", real=True) for i in range(3)]
        xx = self.xx
        # Coordinate-system parameters (C struct members).
        self.params: List[sp.Symbol] = []
        f0F = sp.Function("f0_of_xx0_funcform")(xx[0])
        f1F = sp.Function("f1_of_xx1_funcform")(xx[1])
        if coord == "Cartesian":
            self.eigen = "Cartesian"
            self.scale_funcform = [sp.Integer(1)] * 3
            self.r_of_xx0 = None
            self.f0_closed = None
            self.f1_closed = None
            L = sp.Symbol("grid_physical_size", real=True)
            self.xxmin = [-L, -L, -L]
            self.xxmax = [L, L, L]
            self.xx_to_Cart = list(xx)
        elif coord in ("Spherical", "SinhSpherical"):
            self.eigen = "Spherical"
            if coord == "Spherical":
                RMAX = sp.Symbol("RMAX", real=True)
                self.params = [RMAX]
                r = xx[0]
                self.xxmin = [sp.Integer(0), sp.Integer(0), -PI]
                self.xxmax = [RMAX, PI, PI]
            else:
                AMPL = sp.Symbol("AMPL", real=True)
                SINHW = sp.Symbol("SINHW", real=True)
                self.params = [AMPL, SINHW]
                # Sinhv1 radial map, nrpy/reference_metric.py:787-800.
                r = (
                    AMPL
                    * (sp.exp(xx[0] / SINHW) - sp.exp(-xx[0] / SINHW))
                    / (sp.exp(1 / SINHW) - sp.exp(-1 / SINHW))
                )
                self.xxmin = [sp.Integer(0), sp.Integer(0), -PI]
                self.xxmax = [sp.Integer(1), PI, PI]
            self.r_of_xx0 = r
            self.f0_closed = r
            self.f1_closed = sp.sin(xx[1])
            self.scale_funcform = [sp.diff(f0F, xx[0]), f0F, f0F * f1F]
            th, ph = xx[1], xx[2]
            self.xx_to_Cart = [
                r * sp.sin(th) * sp.cos(ph),
                r * sp.sin(th) * sp.sin(ph),
                r * sp.cos(th),
            ]
        else:
            raise ValueError(f"unsupported coordinate system {coord}")

        self._f0F, self._f1F = f0F, f1F
        self._build_funcform()
        self._to_precompute_symbols()

    # ------------------------------------------------------------------
    def _build_funcform(self) -> None:
        xx = self.xx
        s = self.scale_funcform
        self.ReU = [1 / s[i] for i in range(3)]
        self.ReDD = [[s[i] * s[j] for j in range(3)] for i in range(3)]
        self.ghatDD = T.zeros(2)
        for i in range(3):
            self.ghatDD[i][i] = s[i] ** 2
        self.ghatUU, self.detgammahat = T.symm_inverse3x3(self.ghatDD)
        d = sp.diff
        self.detgammahatdD = [d(self.detgammahat, xx[k]) for k in range(3)]
        self.detgammahatdDD = [[d(self.detgammahatdD[k], xx[l]) for l in range(3)] for k in range(3)]
        self.ReUdD = [[d(self.ReU[i], xx[j]) for j in range(3)] for i in range(3)]
        self.ReUdDD = [[[d(self.ReUdD[i][j], xx[k]) for k in range(3)] for j in range(3)] for i in range(3)]
        self.ReDDdD = [[[d(self.ReDD[i][j], xx[k]) for k in range(3)] for j in range(3)] for i in range(3)]
        self.ReDDdDD = [
            [[[d(self.ReDDdD[i][j][k], xx[l]) for l in range(3)] for k in range(3)] for j in range(3)]
            for i in range(3)
        ]
        self.ghatDDdD = [[[d(self.ghatDD[i][j], xx[k]) for k in range(3)] for j in range(3)] for i in range(3)]
        self.ghatDDdDD = [
            [[[d(self.ghatDDdD[i][j][k], xx[l]) for l in range(3)] for k in range(3)] for j in range(3)]
            for i in range(3)
        ]
        G = T.zeros(3)
        for i, k, l in T.idx(3):
            G[i][k][l] = sum(
                sp.Rational(1, 2)
                * self.ghatUU[i][m]
                * (self.ghatDDdD[m][k][l] + self.ghatDDdD[m][l][k] - self.ghatDDdD[k][l][m])
                for m in range(3)
            )
        self.GammahatUDD = G
        self.GammahatUDDdD = [
            [[[d(G[i][j][k], xx[l]) for l in range(3)] for k in range(3)] for j in range(3)]
            for i in range(3)
        ]

    # ------------------------------------------------------------------
    def _replacement_rules(self, expr) -> Dict:
        rules = {}
        for item in sp.preorder_traversal(expr):
            if isinstance(item, sp.Derivative):
                fname = str(item.args[0].func).replace("_funcform", "")
                var, order = item.args[1]
                vs = str(var)[2:]
                rules[item] = sp.Symbol(fname + "__D" + "D" * (order - 1) + vs * order, real=True)
        return rules

    def _replace(self, expr):
        expr = sp.sympify(expr)
        expr = expr.xreplace(self._replacement_rules(expr))
        expr = expr.xreplace(
            {
                self._f0F: sp.Symbol("f0_of_xx0", real=True),
                self._f1F: sp.Symbol("f1_of_xx1", real=True),
            }
        )
        return expr.xreplace(self._constant_derivs)

    def _to_precompute_symbols(self) -> None:
        # Closed forms of each precompute symbol, and constants for those that
        # do not depend on a coordinate.
        self.precompute_defs: Dict[str, sp.Expr] = {}
        self._constant_derivs: Dict[sp.Symbol, sp.Expr] = {}
        if self.f0_closed is not None:
            x0, x1 = self.xx[0], self.xx[1]
            for n, suffix in enumerate(["", "__D0", "__DD00", "__DDD000"]):
                name = "f0_of_xx0" + suffix
                val = sp.diff(self.f0_closed, x0, n) if n else self.f0_closed
                if x0 not in val.free_symbols:
                    self._constant_derivs[sp.Symbol(name, real=True)] = val
                else:
                    self.precompute_defs[name] = val
            for n, suffix in enumerate(["", "__D1", "__DD11"]):
                name = "f1_of_xx1" + suffix
                val = sp.diff(self.f1_closed, x1, n) if n else self.f1_closed
                self.precompute_defs[name] = val

        def conv(obj):
            if isinstance(obj, list):
                return [conv(o) for o in obj]
            return self._replace(obj)

        for attr in [
            "ReU", "ReDD", "ghatDD", "ghatUU", "detgammahat", "detgammahatdD",
            "detgammahatdDD", "ReUdD", "ReUdDD", "ReDDdD", "ReDDdDD", "ghatDDdD",
            "ghatDDdDD", "GammahatUDD", "GammahatUDDdD",
        ]:
            setattr(self, attr, conv(getattr(self, attr)))
        self.scalefactor = conv(self.scale_funcform)

    # ------------------------------------------------------------------
    def precompute_symbol_names(self) -> List[str]:
        """Names of the 1D precompute arrays, xx0-dependent ones first."""
        return list(self.precompute_defs.keys())


_CACHE: Dict[str, RefMetric] = {}


def reference_metric(coord: str) -> RefMetric:
    """Cached reference metric for a coordinate system."""
    if coord not in _CACHE:
        _CACHE[coord] = RefMetric(coord)
    return _CACHE[coord]
