"""Pins for the CPU oracle's hand-written parts (stencil loops, boundary
conditions, RK4, grids), against what the mathematics fixes:

* flat space has vanishing BSSN right-hand sides and Ricci tensor;
* Brill-Lindquist data satisfy the Hamiltonian constraint, and the discrete H
  converges to zero at 4th order;
* RK4 on dy/dt = lambda y multiplies by 1 + z + z^2/2 + z^3/6 + z^4/24;
* det(gammabar) = det(gammahat) after enforcement, which is idempotent;
* the boundary struct classifies every ghost point once, maps inner points to
  valid interior/outer sources, and keeps scalar parity +1;
* the radiation condition leaves a pure 1/r profile static and propagates
  k/r^3 tails exactly;
* the scalar-wave evolution converges to the exact solution at 4th order.
"""

import numpy as np
import pytest

from nrpy_amd import initial_data as ID
from oracle import oracle as O

SPH, SINH = O.SPHERICAL, O.SINHSPHERICAL


def interior(a, NG):
    return a[:, NG:-NG, NG:-NG, NG:-NG]


def flat(g):
    y = g.gfs()
    y[ID.EVOL_INDEX["alpha"]] = 1.0
    y[ID.EVOL_INDEX["cf"]] = 1.0
    return y


@pytest.mark.parametrize("coord,variant,fd_order,size,N", [(SPH, 0, 4, 5.0, (16, 8, 2)), (SINH, 1, 6, 20.0, (12, 8, 8))])
def test_flat_space_is_static(oracle_lib, coord, variant, fd_order, size, N):
    g = O.OracleGrid(coord, N, fd_order, size, 0.2, variant=variant)
    y = flat(g)
    aux = g.ricci(y)
    r = g.rhs(y, aux, g.params())
    assert np.abs(interior(aux, g.NG)).max() < 1e-10
    assert np.abs(interior(r, g.NG)).max() < 1e-10


def _hamiltonian_rms(N):
    g = O.OracleGrid(SPH, (N, N // 2, 2), 4, 6.0, variant=0)
    y = ID.brill_lindquist(SPH, g.xx, 6.0, 0.2, (0.5, 0.5), (0, 0, 0.5), (0, 0, -0.5))
    g.apply_bcs_outerextrap_and_inner(y)
    c = g.constraints(y, g.ricci(y))
    H = interior(c, g.NG)[0]
    r = g.xx[0][g.NG:-g.NG]
    mask = (r > 2.0) & (r < 4.0)  # away from the punctures and the outer boundary
    return np.sqrt(np.mean(H[:, :, mask] ** 2))


def test_brill_lindquist_hamiltonian_converges_4th_order(oracle_lib):
    e1, e2 = _hamiltonian_rms(32), _hamiltonian_rms(64)
    assert e2 < e1
    assert np.log2(e1 / e2) > 3.3, (e1, e2)


def test_rk4_amplification_factor(oracle_lib):
    """Four substeps with k = lambda * stage reproduce the RK4 stability polynomial."""
    lam, dt = -0.7, 0.3
    n = 5
    y_n = np.full(n, 2.0)
    y_run, k_odd, k_even = np.zeros(n), np.zeros(n), np.zeros(n)
    L = oracle_lib.lib()
    stage_in = [y_n, k_odd, k_even, k_odd]
    stage_out = [k_odd, k_even, k_odd, k_even]
    for s in range(1, 5):
        stage_out[s - 1][:] = lam * stage_in[s - 1]
        L.og_rk_substep(s, n, dt, stage_out[s - 1].ctypes.data, y_n.ctypes.data, y_run.ctypes.data)
    z = lam * dt
    assert np.allclose(y_n, 2.0 * (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24), rtol=1e-15, atol=0)


@pytest.mark.parametrize("coord,variant", [(SPH, 0), (SINH, 1)])
def test_detgamma_enforcement(oracle_lib, coord, variant):
    g = O.OracleGrid(coord, (10, 6, 4 if coord == SINH else 2), 4 if variant == 0 else 6, 5.0, 0.2, variant=variant)
    y = ID.smooth_perturbation(flat(g), g.xx, seed=4, amplitude=5e-2)
    out = g.enforce_detgamma(y)
    h = out[11:17]  # hDD00..hDD22
    m = np.array([[1 + h[0], h[1], h[2]], [h[1], 1 + h[3], h[4]], [h[2], h[4], 1 + h[5]]])
    det = np.linalg.det(np.moveaxis(m, (0, 1), (-2, -1)))
    assert np.abs(det - 1).max() < 1e-13
    assert np.abs(g.enforce_detgamma(out) - out).max() < 1e-14
    assert np.array_equal(np.delete(out, range(11, 17), 0), np.delete(y, range(11, 17), 0))


@pytest.mark.parametrize("coord,N", [(SPH, (12, 6, 2)), (SINH, (10, 8, 6))])
def test_bcstruct_invariants(oracle_lib, coord, N):
    g = O.OracleGrid(coord, N, 4, 5.0, 0.2, variant=0)
    NG, Nx = g.NG, g.Nx
    dst, src, par = g.inner_bc()
    outer = np.concatenate([g.outer_bc(gz, d) for gz in range(NG) for d in range(3)])
    oidx = outer[:, 0].astype(np.int64) + Nx[0] * (outer[:, 1] + Nx[1] * outer[:, 2].astype(np.int64))
    allpts = np.arange(Nx[0] * Nx[1] * Nx[2])
    i0, i1, i2 = allpts % Nx[0], (allpts // Nx[0]) % Nx[1], allpts // (Nx[0] * Nx[1])
    ghost = ~((i0 >= NG) & (i0 < Nx[0] - NG) & (i1 >= NG) & (i1 < Nx[1] - NG) & (i2 >= NG) & (i2 < Nx[2] - NG))
    both = np.concatenate([dst, oidx])
    assert len(both) == len(np.unique(both)) == ghost.sum()
    assert set(both) == set(allpts[ghost])
    # inner sources are interior points or pure outer points, never inner points
    assert not set(src) & set(dst)
    assert np.all(par[:, 0] == 1)
    assert np.all(np.abs(par) == 1)
    # in spherical-like coordinates the pure outer points all lie beyond r_max
    assert np.all(outer[:, 0] >= Nx[0] - NG) and np.all(outer[:, 3] == -1)


def test_radiation_bc_on_1_over_r(oracle_lib):
    """f = f_inf + A/r is a static outgoing solution: with a k/r^3 RHS inside, the ghost
    RHS equals k/r^3 up to FD truncation of d/dr."""
    g = O.OracleGrid(SPH, (48, 4, 2), 4, 10.0, variant=0)
    r = g.xx[0]
    fi = np.zeros(24)
    fi[6] = fi[10] = 1.0  # f_inf of alpha and cf
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
    # one-sided d/dr truncation (~h^4 f^(5)) only; a missing or mis-signed term is O(10x) off
    assert np.abs(ghost - expect).max() < 1e-3 * np.abs(expect).max()


def _wave_error(N, steps):
    """Max error of uu inside r < 5, where the outer extrapolation BC has no influence yet."""
    g = O.OracleGrid(O.CARTESIAN, (N, N, N), 4, 10.0)
    y = ID.wave_spherical_gaussian(g.xx, 0.0, sigma=2.0)
    st = g.new_state(y, nvars=2)
    dt = g.cfl_dt(0.5)
    for _ in range(steps):
        g.wave_mol_step(st, dt)
    exact = ID.wave_spherical_gaussian(g.xx, steps * dt, sigma=2.0)
    X, Y, Z = ID.cartesian_coords(0, g.xx, 0.0, 0.0)
    mask = np.sqrt(X * X + Y * Y + Z * Z) < 5.0
    return np.abs(st["y_n"][0] - exact[0])[mask].max()


@pytest.mark.slow
def test_wave_converges_4th_order(oracle_lib):
    e1 = _wave_error(32, 4)
    e2 = _wave_error(64, 8)
    assert np.log2(e1 / e2) > 3.3, (e1, e2)


def test_cfl_timestep_spherical(oracle_lib):
    """dt = CFL * min(ds): for Spherical the minimum is r_min * dtheta at the first cell."""
    g = O.OracleGrid(SPH, (72, 12, 2), 4, 7.5, variant=0)
    dr, dth = 7.5 / 72, np.pi / 12
    expect = 0.5 * min(dr, (dr / 2) * dth, (dr / 2) * np.sin(dth / 2) * np.pi)
    assert abs(g.cfl_dt(0.5) - expect) < 1e-15
