"""The SymPy transcription against the reference's trusted 30-digit values.

The reference stores, per equation module, expression values with every free
symbol seeded from md5(name) (fixtures extracted into tests/golden/trusted/ by
tests/golden/make_trusted_fixtures.py).  Agreement at relative 1e-24 pins every
term of the algebra the HIP kernels and the oracle are generated from: a
dropped term, wrong sign, index or factor changes the value at the 1e-1..1e-16
level.
"""

import json
import os

import pytest
import sympy as sp

from nrpy_amd.symbolic import bssn, tensors
from oracle import mpf_eval

GOLD = os.path.join(os.path.dirname(__file__), "golden", "trusted")


def trusted(name):
    with open(os.path.join(GOLD, name + ".json")) as f:
        return json.load(f)["values"]


@pytest.fixture(autouse=True)
def _no_symmetry_axes():
    old = tensors.get_symmetry_axes()
    tensors.set_symmetry_axes("")
    yield
    tensors.set_symmetry_axes(old)


def check(exprs, values, keys=None):
    keys = keys or list(values)
    bad = []
    for k in keys:
        v = mpf_eval.evaluate(exprs[k])
        if not mpf_eval.agrees(v, values[k]):
            bad.append((k, str(v)[:20], str(values[k])[:20]))
    assert not bad, bad


def test_rhs_eval_with_ricci_gridfunctions():
    """rhs_eval, SinhSpherical rfm-precompute, RbarDD as gridfunctions, KO on (rhs_eval.py:416-447)."""
    ex = bssn.bssn_rhs_exprs("SinhSpherical", ricci_symbolic=True, enable_KO=True)
    vals = trusted("rhs_eval_RbarSymbolic")
    assert set(vals) == set(ex)
    check(ex, vals)


@pytest.mark.slow
def test_rhs_eval_with_ricci_expanded():
    ex = bssn.bssn_rhs_exprs("SinhSpherical", ricci_symbolic=False, enable_KO=True)
    check(ex, trusted("rhs_eval_RbarExpanded"))


def test_ricci_tensor():
    """Ricci_exprs_k of BSSN_quantities (sorted RbarDD names)."""
    ex = bssn.ricci_exprs("SinhSpherical")
    vals = trusted("BSSN_quantities_SinhSpherical_rfm_precompute")
    names = sorted(ex)
    check({f"Ricci_exprs_{k}": ex[n] for k, n in enumerate(names)}, vals, [f"Ricci_exprs_{k}" for k in range(6)])


def test_gauge_rhss():
    """OnePlusLog + GammaDriving2ndOrder_Covariant without KO (BSSN_gauge_RHSs.py).

    The fixture was generated with stress-energy source terms on
    (BSSN_gauge_RHSs.py:368, enable_T4 = True), which enter B^i through the
    Lambdabar RHS; this vacuum build compares alpha and beta^i only (B^i is
    pinned by the rhs_eval fixtures, which have T4munu off)."""
    ex = bssn.bssn_rhs_exprs("SinhSpherical", ricci_symbolic=True, enable_KO=False)
    vals = trusted("BSSN_gauge_RHSs_SinhSpherical_rfm_precompute")
    check(ex, vals, ["alpha_rhs", "vet_rhsU0", "vet_rhsU1", "vet_rhsU2"])


def test_intermediate_quantities():
    """gammabar^ij, Gammabar^i_jk, Delta^i, Abar^ij, phi derivatives of BSSN_quantities."""
    from nrpy_amd.symbolic.refmetric import reference_metric

    Bq = bssn.BSSNQuantities(reference_metric("SinhSpherical"), ricci_symbolic=False)
    vals = trusted("BSSN_quantities_SinhSpherical_rfm_precompute")

    def flat(name, t, rank):
        out = {}
        for k, idx in enumerate(tensors.idx(rank)):
            v = t
            for i in idx:
                v = v[i]
            out[f"{name}_{k}"] = v
        return out

    ex = {}
    ex.update(flat("gammabarUU", Bq.gammabarUU, 2))
    ex.update(flat("GammabarUDD", Bq.GammabarUDD, 3))
    ex.update(flat("DGammaU", Bq.DGammaU, 1))
    ex.update(flat("AbarUU", Bq.AbarUU, 2))
    ex.update(flat("phi_dBarDD", Bq.phi_dBarDD, 2))
    ex.update(flat("betaU_dDD", Bq.betaU_dDD, 3))
    ex.update(flat("RbarDD", Bq.RbarDD, 2))
    ex["trAbar"] = Bq.trAbar
    ex["detgammabar"] = Bq.detgammabar
    check(ex, vals, sorted(ex))


def test_constraints():
    """H, MU^i, M^2 and the Lambda constraint (BSSN_constraints.py:82-203).

    The fixture was generated with stress-energy sources (BSSN_constraints.py:509,
    "_T4munu"): H += -16 pi rho, MU^i += -8 pi gamma^ij S_j with rho = alpha^2 T^00 and
    S_i = alpha (beta_i T^00 + gamma_ij T^j0) (T4munu.py:20-82, 168-215).  Those source
    terms are restated here only to compare; the product evaluates the vacuum constraints.
    """
    from nrpy_amd.symbolic.refmetric import reference_metric

    ex = bssn.constraint_exprs("SinhSpherical", ricci_symbolic=False)
    Bq = bssn.BSSNQuantities(reference_metric("SinhSpherical"), ricci_symbolic=False)
    T = tensors.declare("T4UU", 2, "sym01", dim=4)
    alpha, PI = sp.Symbol("alpha", real=True), sp.Symbol("PI", real=True)
    gDD = [[Bq.gammabarDD[i][j] / Bq.exp_m4phi for j in range(3)] for i in range(3)]
    gUU = [[Bq.gammabarUU[i][j] * Bq.exp_m4phi for j in range(3)] for i in range(3)]
    betaD = [sum(gDD[i][j] * Bq.betaU[j] for j in range(3)) for i in range(3)]
    SD = [alpha * (betaD[i] * T[0][0] + sum(gDD[i][j] * T[j + 1][0] for j in range(3))) for i in range(3)]
    SU = [sum(gUU[i][j] * SD[j] for j in range(3)) for i in range(3)]
    H = ex["H"] - 16 * PI * alpha**2 * T[0][0]
    MU = [ex[f"MU{i}"] - 8 * PI * SU[i] for i in range(3)]
    Msq = sum(Bq.gammabarDD[i][j] * MU[i] * MU[j] for i in range(3) for j in range(3)) / Bq.exp_m4phi
    vals = trusted("BSSN_constraints_SinhSpherical_rfm_precompute")
    check({"H": H, "MU_0": MU[0], "MU_1": MU[1], "MU_2": MU[2], "Msquared": Msq}, vals,
          ["H", "MU_0", "MU_1", "MU_2", "Msquared"])
    # Lambda-constraint pieces that carry no source terms: Lambdabar^i - Delta^i
    lam = {f"LambdaConstraintU_{i}": Bq.LambdabarU[i] - Bq.DGammaU[i] for i in range(3)}
    lam["LambdaConstraintMagnitude"] = ex["LAMBDA"]
    lam["LambdaConstraintSquared"] = ex["LAMBDA"] ** 2
    check(lam, vals, list(lam))


def test_wave_rhs():
    from nrpy_amd.codegen import specs

    spec = specs.wave_spec()
    ex = {f"{g}_rhs": e for g, e in spec.outputs}
    check(ex, trusted("WaveEquation_RHSs"))


def test_brill_lindquist_formula():
    """psi^4 delta_ij and alpha = psi^-2 of InitialData_Cartesian BrillLindquist, via the
    same formula initial_data.brill_lindquist evaluates (md5 inputs, mpf arithmetic)."""
    x, y, z = sp.symbols("x y z", real=True)
    m1, m2 = sp.symbols("BH1_mass BH2_mass", real=True)
    p1 = sp.symbols("BH1_posn_x BH1_posn_y BH1_posn_z", real=True)
    p2 = sp.symbols("BH2_posn_x BH2_posn_y BH2_posn_z", real=True)
    psi = 1 + m1 / (2 * sp.sqrt((x - p1[0]) ** 2 + (y - p1[1]) ** 2 + (z - p1[2]) ** 2))
    psi += m2 / (2 * sp.sqrt((x - p2[0]) ** 2 + (y - p2[1]) ** 2 + (z - p2[2]) ** 2))
    vals = trusted("InitialData_Cartesian_BrillLindquist")
    check({"gammaDD_0": psi**4, "gammaDD_4": psi**4, "gammaDD_8": psi**4, "alpha": psi**-2,
           "gammaDD_1": sp.Integer(0), "KDD_0": sp.Integer(0)}, vals,
          ["gammaDD_0", "gammaDD_4", "gammaDD_8", "alpha", "gammaDD_1", "KDD_0"])


def test_brill_lindquist_numpy_matches_formula():
    """The numpy initial-data code evaluates psi^-2 as the pinned formula does."""
    import numpy as np

    from nrpy_amd import initial_data as ID

    xx = [np.array([0.3, 1.1]), np.array([0.4, 2.0]), np.array([-1.0, 0.5])]
    y = ID.brill_lindquist(1, xx, 10.0, 0.2, (0.6, 0.4), (0.3, 0.2, 0.5), (-0.3, -0.2, -0.5))
    r, th, ph = 1.1, 2.0, 0.5
    X, Y, Z = r * np.sin(th) * np.cos(ph), r * np.sin(th) * np.sin(ph), r * np.cos(th)
    psi = 1 + 0.6 / (2 * np.sqrt((X - 0.3) ** 2 + (Y - 0.2) ** 2 + (Z - 0.5) ** 2))
    psi += 0.4 / (2 * np.sqrt((X + 0.3) ** 2 + (Y + 0.2) ** 2 + (Z + 0.5) ** 2))
    assert abs(y[ID.EVOL_INDEX["cf"], 1, 1, 1] - psi**-2) < 1e-15
    assert abs(y[ID.EVOL_INDEX["alpha"], 1, 1, 1] - psi**-2) < 1e-15


def test_mutation_is_caught():
    """A planted mistake in the transcription (a flipped sign) fails the trusted comparison."""
    ex = bssn.bssn_rhs_exprs("SinhSpherical", ricci_symbolic=True, enable_KO=True)
    vals = trusted("rhs_eval_RbarSymbolic")
    alpha, trK = sp.symbols("alpha trK", real=True)
    wrong = ex["trK_rhs"] - 2 * sp.Rational(1, 3) * alpha * trK * trK  # 1/3 alpha K^2 -> -1/3
    assert not mpf_eval.agrees(mpf_eval.evaluate(wrong), vals["trK_rhs"])
