"""General initial data on the device (bhah_hip_initial_data_convert_ADM_Cartesian_to_BSSN)
against the oracle's restatement of the reference's ADM -> BSSN reader on the same
non-conformally-flat ADM data (1e-12 of each gridfunction's scale), and Brill-Lindquist
ADM data against the closed-form Brill-Lindquist kernel."""

import numpy as np
import pytest

from nrpy_amd import initial_data as ID
from tests.analytic_fields import adm_data

pytestmark = pytest.mark.gpu


# evolved gridfunctions grouped by tensor: aDD, alpha, betU, cf, hDD, lambdaU, trK, vetU
GROUPS = [range(0, 6), range(6, 7), range(7, 10), range(10, 11), range(11, 17), range(17, 20), range(20, 21),
          range(21, 24)]
LAMBDA = 5


def _err(a, b):
    """Max error of each tensor, relative to that tensor's largest component: a component that
    is zero in exact arithmetic (betU^theta of a radial B^i) is rounding noise on both sides."""
    num = np.abs(a - b).reshape(a.shape[0], -1).max(axis=1)
    den = np.abs(b).reshape(b.shape[0], -1).max(axis=1)
    return np.array([num[list(g)].max() / max(den[list(g)].max(), 1e-300) for g in GROUPS])


@pytest.mark.parametrize("coord,Nxx,fd,sym,KO,variant,axi", [(1, (24, 12, 2), 4, 4, False, 0, True),
                                                            (2, (20, 12, 10), 6, 0, True, 1, False)])
def test_adm_conversion_matches_oracle(hip, oracle_lib, coord, Nxx, fd, sym, KO, variant, axi):
    import torch
    from nrpy_amd.runtime import Patch, PatchConfig

    size = 5.0 if coord == 1 else 10.0
    P = Patch(PatchConfig(coord=coord, Nxx=Nxx, fd_order=fd, symmetry_axes=sym, enable_KO=KO, grid_physical_size=size,
                          SINHW=0.4, eta=2.0))
    O = oracle_lib.OracleGrid(coord, Nxx, fd, size, 0.4, variant=variant)
    adm = adm_data(coord, O.xx, size, 0.4, axisymmetric=axi)
    y_o = O.adm_cartesian_to_bssn(adm)
    y_d = P.new_gfs()
    P.initial_data_from_adm(torch.from_numpy(adm).cuda(), y_d)
    torch.cuda.synchronize()
    er = _err(y_d.cpu().numpy(), y_o)
    print(f"\nADM -> BSSN {Nxx}: per-tensor errors {' '.join(f'{e:.1e}' for e in er)}; lambda scale "
          f"{np.abs(y_o[17:20]).max():.2e}")
    assert np.abs(y_o[17:20]).max() > 1e-3
    # algebraic conversion: 1e-12.  lambda^i comes from finite differences of hDD, which
    # amplify hDD's rounding by sum|w| * invdx (up to 36 here) and 1/ReU (r up to 10): 1e-11.
    assert np.delete(er, LAMBDA).max() < 1e-12, er
    assert er[LAMBDA] < 1e-11, er


def test_brill_lindquist_adm_data_give_the_closed_form(hip, oracle_lib):
    import torch
    from nrpy_amd.runtime import Patch, PatchConfig

    P = Patch(PatchConfig(coord=2, Nxx=(20, 12, 10), fd_order=6, symmetry_axes=0, enable_KO=True,
                          grid_physical_size=10.0, SINHW=0.4))
    xx = P.xx()
    X, Y, Z = ID.cartesian_coords(2, xx, 10.0, 0.4)
    m, p1, p2 = (0.6, 0.4), (0.3, 0.2, 0.5), (-0.3, -0.2, -0.5)
    r1 = np.sqrt((X - p1[0]) ** 2 + (Y - p1[1]) ** 2 + (Z - p1[2]) ** 2)
    r2 = np.sqrt((X - p2[0]) ** 2 + (Y - p2[1]) ** 2 + (Z - p2[2]) ** 2)
    psi = 1 + m[0] / (2 * r1) + m[1] / (2 * r2)
    adm = np.zeros((19,) + X.shape)
    adm[0] = adm[3] = adm[5] = psi**4
    adm[12] = psi**-2
    y_adm = P.new_gfs()
    P.initial_data_from_adm(torch.from_numpy(adm).cuda(), y_adm)
    y_bl = P.new_gfs()
    P.initial_data_brill_lindquist(y_bl, m[0], p1, m[1], p2)
    P.apply_bcs_outerextrap_and_inner(y_bl)
    torch.cuda.synchronize()
    assert (y_adm - y_bl).abs().max().item() < 1e-11
