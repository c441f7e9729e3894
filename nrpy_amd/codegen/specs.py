"""Pointwise-kernel specifications shared by the HIP and C emitters.

A spec is a list of outputs (gridfunction name, SymPy expression) whose free
symbols are classified into: centre values of evolved or auxiliary
gridfunctions, finite-difference derivatives of evolved gridfunctions,
reference-metric precompute arrays, and runtime parameters.

Variants fix the coordinate system, the symmetry axes (derivatives along them
vanish) and whether Kreiss-Oliger dissipation is on; the FD order only
changes stencils, not the algebra.
"""

from __future__ import annotations

import functools
from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import sympy as sp

from ..symbolic import bssn, tensors
from ..symbolic.refmetric import reference_metric
from . import fd

# Reference-metric precompute arrays, fixed slot numbers shared by host setup
# and kernels: rfm0[slot][i0] for xx0 functions, rfm1[slot-4][i1] for xx1.
RFM_SYMBOLS = [
    "f0_of_xx0", "f0_of_xx0__D0", "f0_of_xx0__DD00", "f0_of_xx0__DDD000",
    "f1_of_xx1", "f1_of_xx1__D1", "f1_of_xx1__DD11",
]
PARAM_SYMBOLS = ["eta", "KreissOliger_strength_gauge", "KreissOliger_strength_nongauge", "wavespeed"]

WAVE_GFS = ["uu", "vv"]


@dataclass(frozen=True)
class Variant:
    coord: str
    symmetry_axes: str
    enable_KO: bool

    @property
    def tag(self) -> str:
        sym = self.symmetry_axes or "none"
        return f"{self.coord}_sym{sym}_KO{int(self.enable_KO)}"


# Variants used by the evolution configurations:
#  Spherical, axisymmetric (phi), no KO: the head-on collision family
#  SinhSpherical, full 3D, KO on: the 192x96x96 single-patch workload
VARIANTS = [
    Variant("Spherical", "2", False),
    Variant("SinhSpherical", "", True),
]


@dataclass
class Spec:
    name: str
    outputs: List[Tuple[str, sp.Expr]]
    gfs: List[str]  # evolved gridfunction order (array layout)
    aux: List[str] = field(default_factory=list)
    upwind: List[sp.Expr] | None = None  # control vector betaU[d]
    # staged intermediates (symbol, definition), in dependency order
    defs: List[Tuple[sp.Symbol, sp.Expr]] = field(default_factory=list)

    def symbols(self) -> List[str]:
        names = set()
        for _, e in self.outputs:
            names |= {str(s) for s in e.free_symbols}
        for _, e in self.defs:
            names |= {str(s) for s in e.free_symbols}
        if self.upwind is not None:
            for e in self.upwind:
                names |= {str(s) for s in sp.sympify(e).free_symbols}
        staged = {str(s) for s, _ in self.defs}
        return sorted(names - staged)

    def used_defs(self) -> List[Tuple[sp.Symbol, sp.Expr]]:
        """The staged definitions the outputs actually need (transitively)."""
        dmap = dict(self.defs)
        need = set()
        stack = [s for _, e in self.outputs for s in e.free_symbols if s in dmap]
        while stack:
            s = stack.pop()
            if s in need:
                continue
            need.add(s)
            stack.extend(x for x in dmap[s].free_symbols if x in dmap and x not in need)
        return [(s, e) for s, e in self.defs if s in need]

    def classify(self):
        """Split symbols into (centre gfs, aux gfs, derivs, rfm, params)."""
        centre, aux, derivs, rfm, params = [], [], [], [], []
        for n in self.symbols():
            if n in self.gfs:
                centre.append(n)
            elif n in self.aux:
                aux.append(n)
            elif fd.is_deriv_symbol(n):
                d = fd.parse_deriv(n)
                if d.gf not in self.gfs:
                    raise ValueError(f"derivative of unknown gridfunction {n}")
                derivs.append(d)
            elif n in RFM_SYMBOLS:
                rfm.append(n)
            elif n in PARAM_SYMBOLS:
                params.append(n)
            else:
                raise ValueError(f"unclassified symbol {n} in {self.name}")
        return centre, aux, derivs, rfm, params


def _with_axes(axes: str):
    old = tensors.get_symmetry_axes()
    tensors.set_symmetry_axes(axes)
    return old


@functools.lru_cache(maxsize=None)
def bssn_rhs_spec(v: Variant, staged: bool = False) -> Spec:
    old = _with_axes(v.symmetry_axes)
    stager = bssn.Stager() if staged else None
    try:
        ex = bssn.bssn_rhs_exprs(v.coord, ricci_symbolic=True, enable_KO=v.enable_KO, stager=stager)
        rfm = reference_metric(v.coord)
        vet = tensors.declare("vetU", 1)
        upwind = [vet[i] * rfm.ReU[i] for i in range(3)]
    finally:
        tensors.set_symmetry_axes(old)
    outs = [(_rhs_key_gf(k), e) for k, e in ex.items()]
    outs.sort(key=lambda t: bssn.EVOL_GFS.index(t[0]))
    name = f"bssn_rhs_{v.tag}" + ("_staged" if staged else "")
    return Spec(name, outs, list(bssn.EVOL_GFS), list(bssn.AUXEVOL_GFS), upwind,
                defs=stager.defs if stager else [])


def _rhs_key_gf(key: str) -> str:
    # a_rhsDD01 -> aDD01 ; lambda_rhsU0 -> lambdaU0 ; alpha_rhs -> alpha
    # vet_rhsU1 -> vetU1 ; bet_rhsU2 -> betU2 ; cf_rhs -> cf ; trK_rhs -> trK
    base, _, suffix = key.partition("_rhs")
    return base + suffix


@functools.lru_cache(maxsize=None)
def ricci_spec(v: Variant, staged: bool = False) -> Spec:
    old = _with_axes(v.symmetry_axes)
    stager = bssn.Stager() if staged else None
    try:
        ex = bssn.ricci_exprs(v.coord, stager=stager)
    finally:
        tensors.set_symmetry_axes(old)
    name = f"ricci_{v.tag}" + ("_staged" if staged else "")
    return Spec(name, list(ex.items()), list(bssn.EVOL_GFS), [], defs=stager.defs if stager else [])


@functools.lru_cache(maxsize=None)
def detgamma_spec(v: Variant) -> Spec:
    ex = bssn.detgamma_enforce_exprs(v.coord)
    return Spec(f"detgamma_{v.tag}", list(ex.items()), list(bssn.EVOL_GFS), [])


@functools.lru_cache(maxsize=None)
def initial_lambda_spec(v: Variant) -> Spec:
    old = _with_axes(v.symmetry_axes)
    try:
        ex = bssn.initial_lambda_exprs(v.coord)
    finally:
        tensors.set_symmetry_axes(old)
    return Spec(f"initial_lambda_{v.tag}", list(ex.items()), list(bssn.EVOL_GFS), [])


@functools.lru_cache(maxsize=None)
def constraints_spec(v: Variant) -> Spec:
    old = _with_axes(v.symmetry_axes)
    try:
        ex = bssn.constraint_exprs(v.coord)
    finally:
        tensors.set_symmetry_axes(old)
    return Spec(f"constraints_{v.tag}", list(ex.items()), list(bssn.EVOL_GFS), list(bssn.AUXEVOL_GFS))


@functools.lru_cache(maxsize=None)
def wave_spec() -> Spec:
    uu_dDD = tensors.declare("uu_dDD", 2, "sym01")
    vv = sp.Symbol("vv", real=True)
    c = sp.Symbol("wavespeed", real=True)
    vv_rhs = sum(c * c * uu_dDD[i][i] for i in range(3))
    return Spec("wave_rhs", [("uu", vv), ("vv", vv_rhs)], list(WAVE_GFS), [])
