"""The HIP path against what the mathematics fixes (the same pins as the oracle's, applied
to the library directly through the C-ABI):

* inner-boundary parity reproduces Cartesian scalar, vector and tensor fields;
* the reference-metric arrays equal SymPy derivatives of the SinhSpherical map;
* quadratic extrapolation reproduces quadratics (the wave's Cartesian boundary);
* the discrete Hamiltonian constraint of Brill-Lindquist data converges at 4th / 6th order
  with the fd4 / fd6 SinhSpherical kernel builds;
* a partial checkpoint payload leaves every unlisted point at zero.
"""

import struct

import numpy as np
import pytest
import sympy as sp

from nrpy_amd import initial_data as ID
from tests.analytic_fields import bssn_field

pytestmark = pytest.mark.gpu


def _patch(coord, Nxx, fd_order, size, sinhw, sym, KO):
    from nrpy_amd.runtime import Patch, PatchConfig

    return Patch(PatchConfig(coord=coord, Nxx=Nxx, fd_order=fd_order, symmetry_axes=sym, enable_KO=KO,
                             grid_physical_size=size, SINHW=sinhw, eta=2.0))


@pytest.mark.parametrize("coord,Nxx,fd,size,sym,KO", [(1, (12, 8, 2), 4, 5.0, 4, False),
                                                       (2, (10, 8, 6), 6, 20.0, 0, True)])
def test_inner_bc_reproduces_cartesian_fields(hip, coord, Nxx, fd, size, sym, KO):
    import torch

    P = _patch(coord, Nxx, fd, size, 0.4, sym, KO)
    exact = bssn_field(coord, P.xx(), size, 0.4)
    dst, _, _ = P.inner_bc()
    f = exact.copy()
    f.reshape(24, -1)[:, dst] = 0.0
    fd_ = torch.from_numpy(np.ascontiguousarray(f)).cuda()
    P.apply_bcs_inner_only(fd_)
    got = fd_.cpu().numpy().reshape(24, -1)
    scale = np.abs(exact).reshape(24, -1).max(axis=1)[:, None]
    err = np.abs(got[:, dst] - exact.reshape(24, -1)[:, dst]) / scale
    assert err.max() < 1e-13, err.max()


def test_rfm_arrays_against_sympy(hip):
    P = _patch(2, (24, 12, 6), 6, 64.0, 0.2, 0, True)
    x, A, w = sp.symbols("x A w", real=True)
    r = A * sp.sinh(x / w) / sp.sinh(1 / w)
    x0, x1, _ = P.xx()
    rfm0, rfm1 = P.rfm()
    for k in range(4):
        want = sp.lambdify((x, A, w), sp.diff(r, x, k), "numpy")(x0, 64.0, 0.2)
        assert np.abs(rfm0[k] - want).max() <= 1e-14 * np.abs(want).max(), k
    assert np.allclose(rfm1, [np.sin(x1), np.cos(x1), -np.sin(x1)], rtol=0, atol=1e-15)


def test_wave_extrapolation_reproduces_quadratics(hip):
    import torch
    from nrpy_amd.runtime import CARTESIAN, SYSTEM_WAVE, Patch, PatchConfig

    P = Patch(PatchConfig(system=SYSTEM_WAVE, coord=CARTESIAN, Nxx=(10, 8, 6), fd_order=4, symmetry_axes=0,
                          grid_physical_size=3.0))
    xx = P.xx()
    Z, Y, X = np.meshgrid(xx[2], xx[1], xx[0], indexing="ij")
    q = 1.0 + 0.3 * X - 0.2 * Y + 0.5 * Z + 0.1 * X * X - 0.07 * Y * Z + 0.04 * X * Y + 0.06 * Z * Z
    f = np.stack([q, 2.0 * q])
    NG = P.NG
    g = f.copy()
    g[:, :NG] = g[:, -NG:] = 0.0
    g[:, :, :NG] = g[:, :, -NG:] = 0.0
    g[:, :, :, :NG] = g[:, :, :, -NG:] = 0.0
    gd = torch.from_numpy(np.ascontiguousarray(g)).cuda()
    P.wave_apply_bcs(gd)
    assert np.abs(gd.cpu().numpy() - f).max() < 1e-12 * np.abs(f).max()


def _h_rms(fd_order, N):
    import torch

    P = _patch(2, N, fd_order, 10.0, 0.4, 0, True)
    xx = P.xx()
    y = torch.from_numpy(np.ascontiguousarray(
        ID.brill_lindquist(2, xx, 10.0, 0.4, (0.5, 0.5), (0, 0, 0.5), (0, 0, -0.5)))).cuda()
    P.apply_bcs_outerextrap_and_inner(y)
    aux, diag = P.new_gfs(6), P.new_gfs(6)
    P.Ricci_eval(y, aux)
    P.diagnostics_constraints(aux, y, diag, integrals=False)
    NG = P.NG
    H = diag.cpu().numpy()[0][NG:-NG, NG:-NG, NG:-NG]
    r = ID.radius_of_xx0(2, xx[0], 10.0, 0.4)[NG:-NG]
    m = (r > 1.5) & (r < 4.0)
    return np.sqrt(np.mean(H[:, :, m] ** 2))


@pytest.mark.parametrize("fd_order,lo,hi", [(4, 3.6, 4.8), (6, 5.5, 7.0)])
def test_hamiltonian_convergence_order(hip, fd_order, lo, hi):
    """Brill-Lindquist H (AMPL 10, SINHW 0.4) at 32x16x2 and 64x32x2 through the HIP Ricci and
    constraint kernels: it falls by about 2^fd_order (the oracle gives 4.24 and 6.13)."""
    e1, e2 = _h_rms(fd_order, (32, 16, 2)), _h_rms(fd_order, (64, 32, 2))
    order = np.log2(e1 / e2)
    print(f"\nfd{fd_order}: H rms {e1:.3e} -> {e2:.3e}, order {order:.2f}")
    assert lo < order < hi, (e1, e2, order)


def test_partial_checkpoint_zeroes_unlisted_points(hip):
    import torch

    P = _patch(1, (8, 4, 2), 4, 5.0, 0.2, 4, False)
    y = torch.full((24, P.Nx[2], P.Nx[1], P.Nx[0]), 7.0, dtype=torch.float64, device="cuda")
    idx = [3, 17, 40]
    vals = np.arange(len(idx) * 24, dtype=np.float64) + 1.0
    payload = struct.pack(f"<{1 + len(idx)}i", len(idx), *idx) + vals.tobytes()
    P.restore_payload(payload, y)
    got = y.cpu().numpy().reshape(24, -1)
    want = np.zeros_like(got)
    for e, i in enumerate(idx):
        want[:, i] = vals[e * 24:(e + 1) * 24]
    assert np.array_equal(got, want)
