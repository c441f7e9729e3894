"""BSSN evolution equations in the reference-metric (rescaled) formalism.

A from-scratch SymPy transcription of the reference's equation modules:

* quantities      ``nrpy/equations/general_relativity/BSSN_quantities.py:36-541``
* RHSs            ``nrpy/equations/general_relativity/BSSN_RHSs.py:28-366``
* gauge RHSs      ``nrpy/equations/general_relativity/BSSN_gauge_RHSs.py:33-334``
  (OnePlusLog lapse, GammaDriving2ndOrder_Covariant shift)
* KO dissipation  ``nrpy/infrastructures/BHaH/general_relativity/rhs_eval.py:166-220``
* det(gammabar) enforcement
  ``nrpy/infrastructures/BHaH/general_relativity/enforce_detgammabar_equals_detgammahat.py:84-98``
* constraints     ``nrpy/equations/general_relativity/BSSN_constraints.py:82-157``

Only the choices used on the evolution path are supported: conformal factor
W, det(gammabar)/det(gammahat) = 1, no stress-energy source.  Symbol names
follow the reference so its trusted 30-digit values apply.
"""

from __future__ import annotations

from typing import Dict, List

import sympy as sp

from . import tensors as T
from .refmetric import RefMetric, reference_metric

R = sp.Rational

# Evolved gridfunctions, in the reference's case-insensitive sorted order
# (nrpy/grid.py:230-268).
EVOL_GFS = [
    "aDD00", "aDD01", "aDD02", "aDD11", "aDD12", "aDD22", "alpha",
    "betU0", "betU1", "betU2", "cf",
    "hDD00", "hDD01", "hDD02", "hDD11", "hDD12", "hDD22",
    "lambdaU0", "lambdaU1", "lambdaU2", "trK", "vetU0", "vetU1", "vetU2",
]
AUXEVOL_GFS = ["RbarDD00", "RbarDD01", "RbarDD02", "RbarDD11", "RbarDD12", "RbarDD22"]

# Outer radiation BC wave speeds and asymptotic values (BSSN_quantities.py:62-86).
WAVESPEED = {g: 1.0 for g in EVOL_GFS}
WAVESPEED["alpha"] = 2.0 ** 0.5
F_INFINITY = {g: 0.0 for g in EVOL_GFS}
F_INFINITY["cf"] = 1.0
F_INFINITY["alpha"] = 1.0


class Stager:
    """Name intermediate tensors so generated code evaluates them once, in stages.

    ``st(name, tensor)`` returns the tensor with each distinct non-trivial entry
    replaced by a symbol ``_<name><i..>`` and records (symbol, definition) in
    ``defs``.  Entries holding upwinded or Kreiss-Oliger derivatives stay inline, so
    those still enter the right-hand sides linearly.  ``expand(e)`` substitutes the
    definitions back (for checking against the unstaged expressions).
    """

    def __init__(self) -> None:
        self.defs: List = []
        self._by_expr: Dict = {}

    def __call__(self, name: str, t, _idx: str = ""):
        if isinstance(t, list):
            return [self(name, x, _idx + str(i)) for i, x in enumerate(t)]
        t = sp.sympify(t)
        if t.is_Atom or any(("_dupD" in str(x) or "_dKOD" in str(x)) for x in t.free_symbols):
            return t
        key = sp.srepr(t)
        if key not in self._by_expr:
            sym = sp.Symbol(f"_{name}{_idx}", real=True)
            self._by_expr[key] = sym
            self.defs.append((sym, t))
        return self._by_expr[key]

    def expand(self, e):
        sub = {}
        for sym, d in self.defs:
            sub[sym] = d.xreplace(sub)
        return sp.sympify(e).xreplace(sub)


def _nostage(name, t):
    return t


def parity_type(gf: str) -> int:
    """Inner-BC parity type of a gridfunction (nrpy/grid.py:116-126, 270-321)."""
    if gf[-2:].isdigit() and ("DD" in gf):
        table = {"00": 4, "01": 5, "02": 6, "11": 7, "12": 8, "22": 9}
        return table[gf[-2:]]
    if gf[-1].isdigit() and "U" in gf:
        return int(gf[-1]) + 1
    return 0


class BSSNQuantities:
    """Conformal BSSN tensors built from the rescaled gridfunctions."""

    def __init__(self, rfm: RefMetric, ricci_symbolic: bool, stager=None):
        self.rfm = rfm
        st = stager or _nostage
        self.st = st
        s = sp.Symbol
        self.hDD = T.declare("hDD", 2, "sym01")
        self.aDD = T.declare("aDD", 2, "sym01")
        self.lambdaU = T.declare("lambdaU", 1)
        self.vetU = T.declare("vetU", 1)
        self.betU = T.declare("betU", 1)
        self.trK, self.cf, self.alpha = s("trK", real=True), s("cf", real=True), s("alpha", real=True)
        I3 = range(3)

        # gammabar_ij = h_ij ReDD_ij + ghat_ij ; Abar_ij = a_ij ReDD_ij
        self.gammabarDD = st("gbDD", [[self.hDD[i][j] * rfm.ReDD[i][j] + rfm.ghatDD[i][j] for j in I3] for i in I3])
        self.AbarDD = st("AbDD", [[self.aDD[i][j] * rfm.ReDD[i][j] for j in I3] for i in I3])
        self.LambdabarU = [self.lambdaU[i] * rfm.ReU[i] for i in I3]
        self.betaU = [self.vetU[i] * rfm.ReU[i] for i in I3]
        self.BU = [self.betU[i] * rfm.ReU[i] for i in I3]

        gUU, _ = T.symm_inverse3x3(self.gammabarDD)
        self.gammabarUU = gUU = st("gbUU", gUU)

        hDD_dD = T.declare("hDD_dD", 3, "sym01")
        hDD_dupD = T.declare("hDD_dupD", 3, "sym01")
        hDD_dDD = T.declare("hDD_dDD", 4, "sym01_sym23")
        self.gammabarDD_dD = T.zeros(3)
        self.gammabarDD_dupD = T.zeros(3)
        for i, j, k in T.idx(3):
            base = rfm.ghatDDdD[i][j][k] + self.hDD[i][j] * rfm.ReDDdD[i][j][k]
            self.gammabarDD_dD[i][j][k] = base + hDD_dD[i][j][k] * rfm.ReDD[i][j]
            self.gammabarDD_dupD[i][j][k] = base + hDD_dupD[i][j][k] * rfm.ReDD[i][j]
        self.gammabarDD_dD = st("gbDDdD", self.gammabarDD_dD)

        # Gammabar^i_kl
        G = T.zeros(3)
        dgb = self.gammabarDD_dD
        for i, k, l in T.idx(3):
            G[i][k][l] = sum(R(1, 2) * gUU[i][m] * (dgb[m][k][l] + dgb[m][l][k] - dgb[k][l][m]) for m in I3)
        self.GammabarUDD = G = st("GbUDD", G)

        # det(gammabar) = det(gammahat) (detgbarOverdetghat_equals_one).
        self.detgammabar = rfm.detgammahat
        self.detgammabar_dD = list(rfm.detgammahatdD)
        self.detgammabar_dDD = [list(row) for row in rfm.detgammahatdDD]

        # Abar^{ij}, Abar^i_j, trace
        A = self.AbarDD
        self.AbarUD = st("AbUD", [[sum(gUU[i][k] * A[k][j] for k in I3) for j in I3] for i in I3])
        self.AbarUU = st("AbUU", [[sum(self.AbarUD[i][l] * gUU[j][l] for l in I3) for j in I3] for i in I3]) \
            if stager else [[sum(gUU[i][k] * gUU[j][l] * A[k][l] for k in I3 for l in I3) for j in I3] for i in I3]
        self.trAbar = st("trAb", sum(gUU[k][j] * A[j][k] for k in I3 for j in I3))
        aDD_dD = T.declare("aDD_dD", 3, "sym01")
        aDD_dupD = T.declare("aDD_dupD", 3, "sym01")
        self.AbarDD_dD = T.zeros(3)
        self.AbarDD_dupD = T.zeros(3)
        for i, j, k in T.idx(3):
            self.AbarDD_dupD[i][j][k] = rfm.ReDDdD[i][j][k] * self.aDD[i][j] + rfm.ReDD[i][j] * aDD_dupD[i][j][k]
            self.AbarDD_dD[i][j][k] = rfm.ReDDdD[i][j][k] * self.aDD[i][j] + rfm.ReDD[i][j] * aDD_dD[i][j][k]

        # Delta^i_jk = Gammabar - Gammahat ; Delta^i
        self.DGammaUDD = st("DGUDD", [[[G[i][j][k] - rfm.GammahatUDD[i][j][k] for k in I3] for j in I3] for i in I3])
        self.DGammaU = st("DGU", [sum(gUU[j][k] * self.DGammaUDD[i][j][k] for j in I3 for k in I3) for i in I3])

        if ricci_symbolic:
            self.RbarDD = T.declare("RbarDD", 2, "sym01")
        else:
            self.RbarDD = self._ricci(hDD_dD, hDD_dDD)

        # Shift derivatives (beta^i = vet^i ReU^i)
        vetU_dD = T.declare("vetU_dD", 2)
        vetU_dupD = T.declare("vetU_dupD", 2)
        vetU_dDD = T.declare("vetU_dDD", 3, "sym12")
        self.betaU_dD = T.zeros(2)
        self.betaU_dupD = T.zeros(2)
        self.betaU_dDD = T.zeros(3)
        for i, j in T.idx(2):
            self.betaU_dD[i][j] = vetU_dD[i][j] * rfm.ReU[i] + self.vetU[i] * rfm.ReUdD[i][j]
            self.betaU_dupD[i][j] = vetU_dupD[i][j] * rfm.ReU[i] + self.vetU[i] * rfm.ReUdD[i][j]
            for k in I3:
                self.betaU_dDD[i][j][k] = (
                    vetU_dDD[i][j][k] * rfm.ReU[i]
                    + vetU_dD[i][j] * rfm.ReUdD[i][k]
                    + vetU_dD[i][k] * rfm.ReUdD[i][j]
                    + self.vetU[i] * rfm.ReUdDD[i][j][k]
                )
        self.betaU_dD = st("bUdD", self.betaU_dD)
        self.betaU_dDD = st("bUdDD", self.betaU_dDD)

        # Conformal factor W = e^{-2 phi}
        cf = self.cf
        cf_dD = T.declare("cf_dD", 1)
        cf_dupD = T.declare("cf_dupD", 1)
        cf_dDD = T.declare("cf_dDD", 2, "sym01")
        self.phi_dD = st("phidD", [-cf_dD[i] / (2 * cf) for i in I3])
        self.phi_dupD = [-cf_dupD[i] / (2 * cf) for i in I3]
        self.phi_dDD = [[(-cf_dDD[i][j] + cf_dD[i] * cf_dD[j] / cf) / (2 * cf) for j in I3] for i in I3]
        self.exp_m4phi = cf * cf
        self.phi_dBarD = self.phi_dD
        self.phi_dBarDD = st("phidBDD", [
            [self.phi_dDD[i][j] - sum(G[k][i][j] * self.phi_dD[k] for k in I3) for j in I3] for i in I3
        ])

    def _ricci(self, hDD_dD, hDD_dDD):
        """Conformal Ricci tensor (BSSN_quantities.py:261-449)."""
        rfm = self.rfm
        I3 = range(3)
        gUU = self.gammabarUU
        Gh = rfm.GammahatUDD
        st = self.st
        eps = st("eps", [[self.hDD[i][j] * rfm.ReDD[i][j] for j in I3] for i in I3])
        eps_dD = T.zeros(3)
        for i, j, k in T.idx(3):
            eps_dD[i][j][k] = hDD_dD[i][j][k] * rfm.ReDD[i][j] + self.hDD[i][j] * rfm.ReDDdD[i][j][k]
        eps_dDD = T.zeros(4)
        for i, j, k, l in T.idx(4):
            eps_dDD[i][j][k][l] = (
                hDD_dDD[i][j][k][l] * rfm.ReDD[i][j]
                + hDD_dD[i][j][k] * rfm.ReDDdD[i][j][l]
                + hDD_dD[i][j][l] * rfm.ReDDdD[i][j][k]
                + self.hDD[i][j] * rfm.ReDDdDD[i][j][k][l]
            )
        eps_dD = st("epsdD", eps_dD)
        eps_dDD = st("epsdDD", eps_dDD)
        # \bar{gamma}_{ij;\hat{l}}
        gDhat = T.zeros(3)
        for i, j, l in T.idx(3):
            gDhat[i][j][l] = eps_dD[i][j][l] + sum(
                -Gh[m][i][l] * eps[m][j] - Gh[m][j][l] * eps[i][m] for m in I3
            )
        self.gammabarDD_dHatD = gDhat = st("gDh", gDhat)
        # \bar{gamma}_{ij;\hat{l},k}
        gDhat_dD = T.zeros(4)
        for i, j, l, k in T.idx(4):
            gDhat_dD[i][j][l][k] = eps_dDD[i][j][l][k] + sum(
                -rfm.GammahatUDDdD[m][i][l][k] * eps[m][j]
                - Gh[m][i][l] * eps_dD[m][j][k]
                - rfm.GammahatUDDdD[m][j][l][k] * eps[i][m]
                - Gh[m][j][l] * eps_dD[i][m][k]
                for m in I3
            )
        gDhat_dD = st("gDhdD", gDhat_dD)
        # \bar{gamma}_{ij;\hat{l}\hat{k}}
        gDhatDD = T.zeros(4)
        for i, j, l, k in T.idx(4):
            gDhatDD[i][j][l][k] = gDhat_dD[i][j][l][k] + sum(
                -Gh[m][l][k] * gDhat[i][j][m] - Gh[m][i][k] * gDhat[m][j][l] - Gh[m][j][k] * gDhat[i][m][l]
                for m in I3
            )
        gDhatDD = st("gDhDD", gDhatDD)
        # \hat{D}_j \bar{Lambda}^k
        lambdaU_dD = T.declare("lambdaU_dD", 2)
        LamDhat = T.zeros(2)
        for j, k in T.idx(2):
            LamDhat[k][j] = (
                lambdaU_dD[k][j] * rfm.ReU[k]
                + self.lambdaU[k] * rfm.ReUdD[k][j]
                + sum(Gh[k][m][j] * self.lambdaU[m] * rfm.ReU[m] for m in I3)
            )
        DG = self.DGammaUDD
        gDD = self.gammabarDD
        LamDhat = st("LDh", LamDhat)
        DGammaDDD = st("DGDDD", [[[sum(gDD[i][m] * DG[m][j][k] for m in I3) for k in I3] for j in I3] for i in I3])
        Rb = T.zeros(2)
        for i, j in T.idx(2):
            Rb[i][j] += sum(-R(1, 2) * gUU[k][l] * gDhatDD[i][j][l][k] for k in I3 for l in I3)
            Rb[i][j] += sum(R(1, 2) * (gDD[k][i] * LamDhat[k][j] + gDD[k][j] * LamDhat[k][i]) for k in I3)
            Rb[i][j] += sum(
                R(1, 2) * self.DGammaU[k] * (DGammaDDD[i][j][k] + DGammaDDD[j][i][k]) for k in I3
            )
            Rb[i][j] += sum(
                gUU[k][l]
                * (
                    DG[m][k][i] * DGammaDDD[j][m][l]
                    + DG[m][k][j] * DGammaDDD[i][m][l]
                    + DG[m][i][k] * DGammaDDD[m][j][l]
                )
                for k in I3
                for l in I3
                for m in I3
            )
        return Rb


def ricci_exprs(coord: str, stager=None) -> Dict[str, sp.Expr]:
    """RbarDD{ij} (i<=j) in terms of hDD, its derivatives and lambdaU_dD."""
    Bq = BSSNQuantities(reference_metric(coord), ricci_symbolic=False, stager=stager)
    return {f"RbarDD{i}{j}": Bq.RbarDD[i][j] for i in range(3) for j in range(i, 3)}


def bssn_rhs_exprs(
    coord: str,
    ricci_symbolic: bool = True,
    enable_KO: bool = True,
    stager=None,
) -> Dict[str, sp.Expr]:
    """All 24 BSSN right-hand sides, keyed ``<gf>_rhs`` style names.

    Keys follow the reference's ``BSSN_RHSs_varname_to_expr_dict`` plus the gauge
    terms added in ``rhs_eval.py:129-139`` (``alpha_rhs``, ``vet_rhsU{i}``,
    ``bet_rhsU{i}``), and KO dissipation from ``rhs_eval.py:166-220``.
    """
    rfm = reference_metric(coord)
    Bq = BSSNQuantities(rfm, ricci_symbolic=ricci_symbolic, stager=stager)
    st = Bq.st
    I3 = range(3)
    gDD, gUU, A = Bq.gammabarDD, Bq.gammabarUU, Bq.AbarDD
    alpha, trK, cf = Bq.alpha, Bq.trK, Bq.cf
    betaU, bdD, bdDD = Bq.betaU, Bq.betaU_dD, Bq.betaU_dDD
    detg, detg_dD, detg_dDD = Bq.detgammabar, Bq.detgammabar_dD, Bq.detgammabar_dDD

    # ---- gammabar RHS (BSSN_RHSs.py:79-117)
    gam_rhs = T.zeros(2)
    for i, j in T.idx(2):
        gam_rhs[i][j] = sum(
            betaU[k] * Bq.gammabarDD_dupD[i][j][k] + bdD[k][i] * gDD[k][j] + bdD[k][j] * gDD[i][k] for k in I3
        )
    Dbarbeta = st("Dbb", sum(bdD[k][k] + betaU[k] * detg_dD[k] / (2 * detg) for k in I3))
    for i, j in T.idx(2):
        gam_rhs[i][j] += R(2, 3) * gDD[i][j] * (alpha * Bq.trAbar - Dbarbeta)
    for i, j in T.idx(2):
        gam_rhs[i][j] += -2 * alpha * A[i][j]

    # ---- Abar RHS (BSSN_RHSs.py:119-194)
    A_rhs = T.zeros(2)
    for i, j in T.idx(2):
        A_rhs[i][j] = sum(
            betaU[k] * Bq.AbarDD_dupD[i][j][k] + bdD[k][i] * A[k][j] + bdD[k][j] * A[i][k] for k in I3
        )
    for i, j in T.idx(2):
        A_rhs[i][j] += -R(2, 3) * A[i][j] * Dbarbeta + alpha * A[i][j] * trK
        A_rhs[i][j] += sum(-2 * alpha * A[i][k] * Bq.AbarUD[k][j] for k in I3)
    alpha_dD = T.declare("alpha_dD", 1)
    alpha_dDD = T.declare("alpha_dDD", 2, "sym01")
    G = Bq.GammabarUDD
    alpha_dBarDD = st("adBDD", [[alpha_dDD[i][j] - sum(G[k][i][j] * alpha_dD[k] for k in I3) for j in I3] for i in I3])
    phi_dD, phi_dBarDD = Bq.phi_dBarD, Bq.phi_dBarDD
    curly = T.zeros(2)
    for i, j in T.idx(2):
        curly[i][j] = (
            -2 * alpha * phi_dBarDD[i][j]
            + 4 * alpha * phi_dD[i] * phi_dD[j]
            + 2 * alpha_dD[i] * phi_dD[j]
            + 2 * alpha_dD[j] * phi_dD[i]
            - alpha_dBarDD[i][j]
            + alpha * Bq.RbarDD[i][j]
        )
    curly = st("curly", curly)
    curly_tr = st("curlytr", sum(gUU[i][j] * curly[i][j] for i in I3 for j in I3))
    for i, j in T.idx(2):
        A_rhs[i][j] += Bq.exp_m4phi * (curly[i][j] - R(1, 3) * gDD[i][j] * curly_tr)

    # ---- cf RHS, W form (BSSN_RHSs.py:196-216)
    cf_rhs = R(1, 6) * (Dbarbeta - alpha * trK)
    for k in I3:
        cf_rhs += betaU[k] * Bq.phi_dupD[k]
    cf_rhs *= -2 * cf

    # ---- trK RHS (BSSN_RHSs.py:218-241)
    trK_dupD = T.declare("trK_dupD", 1)
    trK_rhs = R(1, 3) * alpha * trK * trK
    for i in I3:
        trK_rhs += betaU[i] * trK_dupD[i]
    for i, j in T.idx(2):
        trK_rhs += -Bq.exp_m4phi * gUU[i][j] * (alpha_dBarDD[i][j] + 2 * alpha_dD[j] * phi_dD[i])
    for i, j in T.idx(2):
        trK_rhs += alpha * A[i][j] * Bq.AbarUU[i][j]

    # ---- Lambdabar RHS (BSSN_RHSs.py:243-331)
    lambdaU_dupD = T.declare("lambdaU_dupD", 2)
    LambdabarU_dupD = [
        [lambdaU_dupD[i][j] * rfm.ReU[i] + Bq.lambdaU[i] * rfm.ReUdD[i][j] for j in I3] for i in I3
    ]
    Lam_rhs = [
        sum(betaU[k] * LambdabarU_dupD[i][k] - bdD[i][k] * Bq.LambdabarU[k] for k in I3) for i in I3
    ]
    Gh, GhdD = rfm.GammahatUDD, rfm.GammahatUDDdD
    for i in I3:
        acc = 0
        for j, k in T.idx(2):
            t2a = bdDD[i][k][j]
            t2b = sum(
                GhdD[i][m][k][j] * betaU[m]
                + Gh[i][m][k] * bdD[m][j]
                + Gh[i][m][j] * bdD[m][k]
                - Gh[m][k][j] * bdD[i][m]
                for m in I3
            )
            t2c = sum(
                (Gh[i][d][j] * Gh[d][m][k] - Gh[d][k][j] * Gh[i][m][d]) * betaU[m] for m in I3 for d in I3
            )
            acc += gUU[j][k] * (t2a + t2b + t2c)
        Lam_rhs[i] += acc
    for i in I3:
        Lam_rhs[i] += R(2, 3) * Bq.DGammaU[i] * Dbarbeta
    Dbb_dBarD = [0, 0, 0]
    for k, m in T.idx(2):
        Dbb_dBarD[m] += (
            bdDD[k][k][m]
            + (bdD[k][m] * detg_dD[k] + betaU[k] * detg_dDD[k][m]) / (2 * detg)
            - betaU[k] * detg_dD[k] * detg_dD[m] / (2 * detg * detg)
        )
    for i, m in T.idx(2):
        Lam_rhs[i] += R(1, 3) * gUU[i][m] * Dbb_dBarD[m]
    for i, j in T.idx(2):
        Lam_rhs[i] += -2 * Bq.AbarUU[i][j] * (alpha_dD[j] - 6 * alpha * phi_dD[j])
    for i, j, k in T.idx(3):
        Lam_rhs[i] += 2 * alpha * Bq.AbarUU[j][k] * Bq.DGammaUDD[i][j][k]
    trK_dD = T.declare("trK_dD", 1)
    for i, j in T.idx(2):
        Lam_rhs[i] += -R(4, 3) * alpha * gUU[i][j] * trK_dD[j]

    # ---- gauge (BSSN_gauge_RHSs.py:99-108, 169-262)
    alpha_dupD = T.declare("alpha_dupD", 1)
    alpha_rhs = -2 * alpha * trK + sum(betaU[i] * alpha_dupD[i] for i in I3)
    BU = Bq.BU
    beta_rhs = [0, 0, 0]
    for i, j in T.idx(2):
        beta_rhs[i] += betaU[j] * Bq.betaU_dupD[i][j]
    for i, j, m in T.idx(3):
        beta_rhs[i] += betaU[j] * G[i][m][j] * betaU[m]
    for i in I3:
        beta_rhs[i] += BU[i]
    betU_dupD = T.declare("betU_dupD", 2)
    BU_dupD = [[betU_dupD[i][j] * rfm.ReU[i] + Bq.betU[i] * rfm.ReUdD[i][j] for j in I3] for i in I3]
    B_rhs = [0, 0, 0]
    for i, j in T.idx(2):
        B_rhs[i] += betaU[j] * BU_dupD[i][j]
    for i, j, m in T.idx(3):
        B_rhs[i] += betaU[j] * G[i][m][j] * BU[m]
    for i in I3:
        B_rhs[i] += R(3, 4) * Lam_rhs[i]
    for i, j in T.idx(2):
        B_rhs[i] += -R(3, 4) * betaU[j] * LambdabarU_dupD[i][j]
    for i, j, m in T.idx(3):
        B_rhs[i] += -R(3, 4) * betaU[j] * G[i][m][j] * Bq.LambdabarU[m]
    eta = sp.Symbol("eta", real=True)
    for i in I3:
        B_rhs[i] += -eta * BU[i]

    # ---- rescale (BSSN_RHSs.py:349-359, BSSN_gauge_RHSs.py:319-323)
    out: Dict[str, sp.Expr] = {}
    out["cf_rhs"] = cf_rhs
    out["trK_rhs"] = trK_rhs
    for i in I3:
        out[f"lambda_rhsU{i}"] = Lam_rhs[i] / rfm.ReU[i]
        for j in range(i, 3):
            out[f"a_rhsDD{i}{j}"] = A_rhs[i][j] / rfm.ReDD[i][j]
            out[f"h_rhsDD{i}{j}"] = gam_rhs[i][j] / rfm.ReDD[i][j]
    out["alpha_rhs"] = alpha_rhs
    for i in I3:
        out[f"vet_rhsU{i}"] = beta_rhs[i] / rfm.ReU[i]
        out[f"bet_rhsU{i}"] = B_rhs[i] / rfm.ReU[i]

    if enable_KO:
        eps_g = sp.Symbol("KreissOliger_strength_gauge", real=True)
        eps_n = sp.Symbol("KreissOliger_strength_nongauge", real=True)
        a_KO = T.declare("alpha_dKOD", 1)
        cf_KO = T.declare("cf_dKOD", 1)
        trK_KO = T.declare("trK_dKOD", 1)
        betU_KO = T.declare("betU_dKOD", 2)
        vetU_KO = T.declare("vetU_dKOD", 2)
        lam_KO = T.declare("lambdaU_dKOD", 2)
        aDD_KO = T.declare("aDD_dKOD", 3, "sym01")
        hDD_KO = T.declare("hDD_dKOD", 3, "sym01")
        for k in I3:
            out["alpha_rhs"] += eps_g * a_KO[k] * rfm.ReU[k]
            out["cf_rhs"] += eps_n * cf_KO[k] * rfm.ReU[k]
            out["trK_rhs"] += eps_n * trK_KO[k] * rfm.ReU[k]
            for i in I3:
                out[f"bet_rhsU{i}"] += eps_g * betU_KO[i][k] * rfm.ReU[k]
                out[f"vet_rhsU{i}"] += eps_g * vetU_KO[i][k] * rfm.ReU[k]
                out[f"lambda_rhsU{i}"] += eps_n * lam_KO[i][k] * rfm.ReU[k]
                for j in range(i, 3):
                    out[f"a_rhsDD{i}{j}"] += eps_n * aDD_KO[i][j][k] * rfm.ReU[k]
                    out[f"h_rhsDD{i}{j}"] += eps_n * hDD_KO[i][j][k] * rfm.ReU[k]
    return dict(sorted(out.items()))


def rhs_key_to_gf(key: str) -> str:
    """Map an RHS key such as ``a_rhsDD01`` to its gridfunction ``aDD01``."""
    base = key.replace("_rhs", "")
    return base


def initial_lambda_exprs(coord: str) -> Dict[str, sp.Expr]:
    """lambda^i of converted initial data: Lambdabar^i = gammabar^{jk} (Gammabar^i_jk - Gammahat^i_jk)
    from finite-difference derivatives of hDD, rescaled by ReU^i
    (ADM_Initial_Data_Reader__BSSN_Converter.py:462-580, math at :506-523)."""
    rfm = reference_metric(coord)
    Bq = BSSNQuantities(rfm, ricci_symbolic=True)
    return {f"lambdaU{i}": Bq.DGammaU[i] / rfm.ReU[i] for i in range(3)}


def detgamma_enforce_exprs(coord: str) -> Dict[str, sp.Expr]:
    """h'_ij = (|det ghat|/det gbar)^{1/3} (ghat_ij/ReDD_ij + h_ij) - ghat_ij/ReDD_ij."""
    rfm = reference_metric(coord)
    hDD = T.declare("hDD", 2, "sym01")
    gbar = [[hDD[i][j] * rfm.ReDD[i][j] + rfm.ghatDD[i][j] for j in range(3)] for i in range(3)]
    _, detgbar = T.symm_inverse3x3(gbar)
    # nrpyAbs stays an opaque function so SymPy does not split |a*b|**(1/3)
    # into per-factor powers; it prints as fabs (custom_c_codegen_functions.py).
    nrpyAbs = sp.Function("nrpyAbs")
    fac = (nrpyAbs(rfm.detgammahat) / detgbar) ** R(1, 3)
    out = {}
    for i in range(3):
        for j in range(i, 3):
            q = rfm.ghatDD[i][j] / rfm.ReDD[i][j]
            out[f"hDD{i}{j}"] = fac * (q + hDD[i][j]) - q
    return out


def constraint_exprs(coord: str, ricci_symbolic: bool = True) -> Dict[str, sp.Expr]:
    """Hamiltonian constraint H, momentum constraint MU^i, M^2 and the magnitude of the
    conformal-connection constraint |Lambdabar^i - Delta^i|.

    Follows ``BSSN_constraints.py:82-203``; with ``ricci_symbolic`` RbarDD is read
    from the auxiliary Ricci gridfunctions, otherwise expanded from hDD.
    """
    rfm = reference_metric(coord)
    Bq = BSSNQuantities(rfm, ricci_symbolic=ricci_symbolic)
    I3 = range(3)
    gUU = Bq.gammabarUU
    H = R(2, 3) * Bq.trK**2
    H += sum(-Bq.AbarDD[i][j] * Bq.AbarUU[i][j] for i in I3 for j in I3)
    Rtr = sum(gUU[i][j] * Bq.RbarDD[i][j] for i in I3 for j in I3)
    p2 = sum(gUU[i][j] * Bq.phi_dBarD[i] * Bq.phi_dBarD[j] for i in I3 for j in I3)
    pDD = sum(gUU[i][j] * Bq.phi_dBarDD[i][j] for i in I3 for j in I3)
    H += Bq.exp_m4phi * (Rtr - 8 * (p2 + pDD))
    MU = [0, 0, 0]
    trK_dD = T.declare("trK_dD", 1)
    for i, j in T.idx(2):
        MU[i] += 6 * Bq.AbarUU[i][j] * Bq.phi_dD[j]
        MU[i] += -R(2, 3) * gUU[i][j] * trK_dD[j]
    G = Bq.GammabarUDD
    AdB = T.zeros(3)
    for i, j, k in T.idx(3):
        AdB[i][j][k] = Bq.AbarDD_dD[i][j][k] + sum(
            -G[l][k][i] * Bq.AbarDD[l][j] - G[l][k][j] * Bq.AbarDD[i][l] for l in I3
        )
    for i, j, k, l in T.idx(4):
        MU[i] += gUU[i][k] * gUU[j][l] * AdB[k][l][j]
    MU = [m * Bq.exp_m4phi for m in MU]
    # M^2 = gammabar_ij MU^i MU^j / e^{-4 phi} (BSSN_constraints.py:177-182)
    Msq = sum(Bq.gammabarDD[i][j] * MU[i] * MU[j] for i in I3 for j in I3) / Bq.exp_m4phi
    # |Lambdabar^i - Delta^i| in the conformal metric (BSSN_constraints.py:190-203)
    LamU = [Bq.LambdabarU[i] - Bq.DGammaU[i] for i in I3]
    LamSq = sum(Bq.gammabarDD[i][j] * LamU[i] * LamU[j] for i in I3 for j in I3)
    return {"H": H, "MU0": MU[0], "MU1": MU[1], "MU2": MU[2], "MSQUARED": Msq, "LAMBDA": sp.sqrt(LamSq)}
