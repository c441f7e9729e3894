"""More pins for the oracle, from what the mathematics fixes:

* inner-boundary parity: Cartesian scalar, vector and tensor fields written in the
  reference's rescaled spherical-like components survive the inner boundary condition
  exactly at every inner ghost point (this pins the ten unit-vector dot products and the
  parity type of each of the 24 gridfunctions, which the fields take from their names);
* the SinhSpherical map: f0 = r(xx0) and its first three derivatives, sin(xx1) and its two,
  r(xx0) of the initial data, and the CFL timestep, against SymPy derivatives of
  r = AMPL sinh(xx0 / SINHW) / sinh(1 / SINHW) (nrpy/reference_metric.py:1217-1270);
* quadratic extrapolation reproduces quadratics (apply_bcs_outerextrap_and_inner__openmp.c:8-34);
* the radiation condition on SinhSpherical (r(xx0) and dxx0/dr inside it) keeps
  f_inf + A / r static;
* the discrete Hamiltonian constraint of Brill-Lindquist data converges at 4th order with
  4th-order stencils and at 6th order with 6th-order stencils on SinhSpherical;
* the volume integrals' whole-domain volume equals the midpoint rule's closed form.
"""

import numpy as np
import pytest
import sympy as sp

from nrpy_amd import initial_data as ID
from oracle import oracle as O
from tests.analytic_fields import bssn_field

SPH, SINH = O.SPHERICAL, O.SINHSPHERICAL


@pytest.mark.parametrize("coord,N,size", [(SPH, (12, 8, 2), 5.0), (SINH, (10, 8, 6), 20.0), (SPH, (8, 6, 8), 3.0)])
def test_inner_bc_reproduces_cartesian_fields(oracle_lib, coord, N, size):
    g = O.OracleGrid(coord, N, 4, size, 0.4, variant=0)
    exact = bssn_field(coord, g.xx, size, 0.4)
    dst, src, par = g.inner_bc()
    assert len(dst) > 0
    f = np.ascontiguousarray(exact.copy())
    f.reshape(24, -1)[:, dst] = 0.0
    g.apply_bcs_inner_only(f)
    scale = np.abs(exact).reshape(24, -1).max(axis=1)[:, None]
    err = np.abs(f.reshape(24, -1)[:, dst] - exact.reshape(24, -1)[:, dst]) / scale
    assert err.max() < 1e-13, err.max()
    # the vector types and the off-diagonal tensor types take both signs somewhere on these
    # grids (the diagonal tensor types are squares of dot products, always +1)
    assert all(set(np.unique(par[:, t])) == {-1, 1} for t in (1, 2, 3, 5, 6, 8))
    assert all(np.all(par[:, t] == 1) for t in (0, 4, 7, 9))


def test_inner_bc_fails_with_a_swapped_parity(oracle_lib):
    """Planted mistake: theta and phi vector parities swapped in the gridfunction table."""
    g = O.OracleGrid(SPH, (12, 8, 2), 4, 5.0, variant=0)
    exact = bssn_field(SPH, g.xx, 5.0, 0.2)
    dst, _, _ = g.inner_bc()
    f = np.ascontiguousarray(exact.copy())
    f.reshape(24, -1)[:, dst] = 0.0
    wrong = list(O.PARITY)
    wrong[21], wrong[22] = wrong[22], wrong[21]  # vetU0 (r) <-> vetU1 (theta)
    g.apply_bcs_inner_only(f, parity=wrong)
    assert np.abs(f.reshape(24, -1)[:, dst] - exact.reshape(24, -1)[:, dst]).max() > 1e-2


def _sinh_map():
    x, A, w = sp.symbols("x A w", real=True)
    r = A * sp.sinh(x / w) / sp.sinh(1 / w)
    return [sp.lambdify((x, A, w), sp.diff(r, x, k), "numpy") for k in range(4)]


@pytest.mark.parametrize("A,w", [(64.0, 0.2), (10.0, 0.4)])
def test_sinhspherical_map_against_sympy(oracle_lib, A, w):
    g = O.OracleGrid(SINH, (24, 12, 6), 6, A, w, variant=1)
    f = _sinh_map()
    x0, x1 = g.xx[0], g.xx[1]
    rfm0 = g.rfm0.reshape(4, -1)
    for k in range(4):
        want = f[k](x0, A, w)
        assert np.abs(rfm0[k] - want).max() <= 1e-14 * np.abs(want).max(), k
    rfm1 = g.rfm1.reshape(3, -1)
    assert np.allclose(rfm1, [np.sin(x1), np.cos(x1), -np.sin(x1)], rtol=0, atol=1e-15)
    r = ID.radius_of_xx0(SINH, x0, A, w)
    assert np.abs(r - f[0](x0, A, w)).max() <= 1e-14 * np.abs(r).max()
    # CFL: dt = CFL min over all points of (|dxx0 r'|, |dxx1 r|, |dxx2 r sin(xx1)|)
    d = g.g.dxx
    R, T = np.meshgrid(x0, x1, indexing="ij")
    ds = np.minimum(np.abs(d[0] * f[1](R, A, w)), np.minimum(np.abs(d[1] * f[0](R, A, w)),
                                                               np.abs(d[2] * f[0](R, A, w) * np.sin(T))))
    dt = 0.5 * ds.min()
    assert abs(g.cfl_dt(0.5) - dt) <= 1e-14 * dt


def test_extrapolation_reproduces_quadratics(oracle_lib):
    # Cartesian wave grid: every ghost point of a quadratic polynomial in (x, y, z)
    g = O.OracleGrid(O.CARTESIAN, (10, 8, 6), 4, 3.0)
    Z, Y, X = np.meshgrid(g.xx[2], g.xx[1], g.xx[0], indexing="ij")
    q = 1.0 + 0.3 * X - 0.2 * Y + 0.5 * Z + 0.1 * X * X - 0.07 * Y * Z + 0.04 * X * Y + 0.06 * Z * Z
    f = np.ascontiguousarray(np.stack([q, 2.0 * q]))
    NG = g.NG
    f2 = f.copy()
    f2[:, :NG] = f2[:, -NG:] = 0.0
    f2[:, :, :NG] = f2[:, :, -NG:] = 0.0
    f2[:, :, :, :NG] = f2[:, :, :, -NG:] = 0.0
    g.wave_apply_bcs(f2)
    assert np.abs(f2 - f).max() < 1e-12 * np.abs(f).max()
    # curvilinear: a scalar quadratic in xx0 times a function of the direction is exact at
    # the outer (r > r_max) ghost points, corners filled by parity from them included
    for coord, size in ((SPH, 5.0), (SINH, 10.0)):
        g = O.OracleGrid(coord, (12, 8, 6), 4, size, 0.4, variant=0)
        P, T, R0 = np.meshgrid(g.xx[2], g.xx[1], g.xx[0], indexing="ij")
        s = (0.5 + 0.3 * R0 - 0.8 * R0 * R0) * (1.0 + 0.3 * np.cos(T) + 0.2 * np.sin(T) * np.cos(P))
        f = np.ascontiguousarray(np.repeat(s[None], 24, axis=0))
        f2 = f.copy()
        NG = g.NG
        f2[:, :, :, -NG:] = 0.0
        g.apply_bcs_outerextrap_and_inner(f2, parity=[0] * 24)
        assert np.abs(f2[:, :, :, -NG:] - f[:, :, :, -NG:]).max() < 1e-12 * np.abs(f).max()


def test_radiation_bc_on_1_over_r_sinhspherical(oracle_lib):
    """f = f_inf + A/r is static and outgoing: with RHS C/r^3 inside, the ghost RHS is C/r^3 up
    to the one-sided d/dr truncation; r(xx0) and dxx0/dr enter every term."""
    g = O.OracleGrid(SINH, (64, 4, 2), 4, 10.0, 0.4, variant=0)
    r = ID.radius_of_xx0(SINH, g.xx[0], 10.0, 0.4)
    fi = np.zeros(24)
    fi[6] = fi[10] = 1.0
    y = g.gfs()
    for gf in range(24):
        y[gf] = fi[gf] + 3.0 / r[None, None, :]
    C = 2.0
    rhs = g.gfs()
    rhs[:] = C / r[None, None, :] ** 3
    NG = g.NG
    rhs[:, :, :, -NG:] = 0.0
    g.apply_bcs_outerradiation_and_inner(y, rhs)
    ghost = rhs[:, NG:-NG, NG:-NG, -NG:]
    expect = C / r[-NG:] ** 3
    assert np.abs(ghost - expect).max() < 1e-3 * np.abs(expect).max()


def _h_rms(N, fd_order):
    g = O.OracleGrid(SINH, N, fd_order, 10.0, 0.4, variant=1)
    y = ID.brill_lindquist(SINH, g.xx, 10.0, 0.4, (0.5, 0.5), (0, 0, 0.5), (0, 0, -0.5))
    g.apply_bcs_outerextrap_and_inner(y)
    c = g.constraints(y, g.ricci(y))
    NG = g.NG
    H = c[0][NG:-NG, NG:-NG, NG:-NG]
    r = ID.radius_of_xx0(SINH, g.xx[0], 10.0, 0.4)[NG:-NG]
    m = (r > 1.5) & (r < 4.0)  # away from the punctures and the outer boundary
    return np.sqrt(np.mean(H[:, :, m] ** 2))


@pytest.mark.parametrize("fd_order,lo,hi", [(4, 3.6, 4.8), (6, 5.5, 7.0)])
def test_sinhspherical_hamiltonian_convergence_order(oracle_lib, fd_order, lo, hi):
    """Brill-Lindquist H on SinhSpherical (AMPL 10, SINHW 0.4) at 32x16x2 and 64x32x2: the
    discrete H falls by 2^fd_order (measured: 4.24 and 6.13)."""
    e1, e2 = _h_rms((32, 16, 2), fd_order), _h_rms((64, 32, 2), fd_order)
    order = np.log2(e1 / e2)
    assert lo < order < hi, (e1, e2, order)


def test_volume_integrals_whole_domain_closed_form(oracle_lib):
    N0, N1, N2, R = 24, 12, 2, 7.5
    g = O.OracleGrid(SPH, (N0, N1, N2), 4, R, variant=0)
    diag = g.gfs(6)
    diag[0] = 1.0  # H = 1: the integral of H^2 is the volume
    vi = g.volume_integrals(diag)
    dr, dth = R / N0, np.pi / N1
    vol = dr**3 * (N0**3 / 3 - N0 / 12) * dth / np.sin(dth / 2) * 2 * np.pi
    assert abs(vi["whole_domain"]["volume"] - vol) < 1e-13 * vol
    assert abs(vi["whole_domain"]["H"]["integral"] - vol) < 1e-13 * vol
    assert abs(vi["whole_domain"]["H"]["RMS"] - 1.0) < 1e-14
    assert vi["sphere_R_gt_8"]["volume"] == 0.0 and vi["sphere_R_gt_8"]["H"]["RMS"] == 0.0
    # a grid reaching past r = 8: the outer recipe is the shells with r > 8 exactly
    g = O.OracleGrid(SPH, (20, 12, 2), 4, 10.0, variant=0)
    diag = g.gfs(6)
    diag[4] = 2.0
    vi = g.volume_integrals(diag)
    r = g.xx[0][g.NG:-g.NG]
    dr = 10.0 / 20
    shells = (r[r > 8] ** 2).sum() * dr * (np.pi / 12) / np.sin(np.pi / 24) * 2 * np.pi / 12 * 12
    assert abs(vi["sphere_R_gt_8"]["volume"] - shells) < 1e-13 * shells
    assert abs(vi["sphere_R_gt_8"]["MSQUARED"]["integral"] - 4 * shells) < 1e-13 * shells


@pytest.mark.parametrize("coord,N,fd,variant,axi", [(SPH, (16, 8, 2), 4, 0, True), (SINH, (14, 8, 6), 6, 1, False)])
def test_adm_to_bssn_conversion_round_trip_and_lambda(oracle_lib, coord, N, fd, variant, axi):
    """Non-conformally-flat ADM data through the oracle's conversion: the ADM fields rebuilt
    from the BSSN variables equal the input at every point except the extrapolated outer
    ghosts (inner ghosts hold parity copies of the same Cartesian point); det(gammabar) =
    det(gammahat); and lambda^i, set from finite differences of h, makes the conformal-
    connection constraint |Lambdabar - Delta| vanish at interior points."""
    from tests.analytic_fields import adm_data, adm_from_bssn

    size = 5.0 if coord == SPH else 10.0
    g = O.OracleGrid(coord, N, fd, size, 0.4, variant=variant)
    adm = adm_data(coord, g.xx, size, 0.4, axisymmetric=axi)
    y = g.adm_cartesian_to_bssn(adm)
    gam, K, beta = adm_from_bssn(g, y)
    NG = g.NG
    keep = np.ones(y.shape[1:], bool)
    keep[:, :, -NG:] = False  # outer (r > r_max) ghosts are extrapolated, not converted
    sym = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for k, (i, j) in enumerate(sym):
        assert np.abs(gam[i, j] - adm[k])[keep].max() < 1e-12 * np.abs(adm[0:6]).max(), ("gamma", i, j)
        assert np.abs(K[i, j] - adm[6 + k])[keep].max() < 1e-12 * np.abs(adm[6:12]).max(), ("K", i, j)
    for i in range(3):
        assert np.abs(beta[i] - adm[13 + i])[keep].max() < 1e-12 * np.abs(adm[13:16]).max()
    c = g.constraints(y, g.ricci(y))
    lam = np.abs(y[17:20, NG:-NG, NG:-NG, NG:-NG]).max()
    assert lam > 1e-3  # the data are not conformally flat: lambda^i is not zero
    # the final extrapolation rewrites h at the outer ghosts after lambda^i was formed, so
    # points whose stencils reach them are left out (as in the reference's order)
    assert np.abs(c[5][NG:-NG, NG:-NG, NG:-2 * NG]).max() < 1e-12 * lam


def test_adm_to_bssn_of_brill_lindquist_is_the_closed_form(oracle_lib):
    """Conformally flat Brill-Lindquist ADM data convert to h = a = 0, W = alpha = psi^-2, K = 0
    and lambda^i = 0 (what initial_data.brill_lindquist writes directly)."""
    size = 5.0
    g = O.OracleGrid(SPH, (16, 8, 2), 4, size, variant=0)
    P, T, R0 = np.meshgrid(g.xx[2], g.xx[1], g.xx[0], indexing="ij")
    X, Y, Z = R0 * np.sin(T) * np.cos(P), R0 * np.sin(T) * np.sin(P), R0 * np.cos(T)
    psi = 1.0 + 0.5 / (2 * np.sqrt(X**2 + Y**2 + (Z - 0.5) ** 2)) + 0.5 / (2 * np.sqrt(X**2 + Y**2 + (Z + 0.5) ** 2))
    adm = np.zeros((19,) + X.shape)
    adm[0] = adm[3] = adm[5] = psi**4
    adm[12] = psi**-2
    y = g.adm_cartesian_to_bssn(adm)
    ref = ID.brill_lindquist(SPH, g.xx, size, 0.2, (0.5, 0.5), (0, 0, 0.5), (0, 0, -0.5))
    g.apply_bcs_outerextrap_and_inner(ref)
    assert np.abs(y - ref).max() < 1e-12
