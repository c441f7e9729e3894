"""bhah_hip_diagnostics_constraints_noalloc (the entry point that allocates nothing) through
the C-ABI against bhah_hip_diagnostics_constraints on the same inputs: the diagnostic
gridfunctions and the volume integrals bitwise equal, and the integrals against a numpy
reduction of the oracle's constraints (1e-12).  Two calls in a row share the GPU's scratch
array, the second on a smaller grid; a third repeats the first grid after it."""

import ctypes

import numpy as np
import pytest

from nrpy_amd import initial_data as ID

pytestmark = pytest.mark.gpu

GRIDS = [(40, 20, 24), (20, 10, 12), (40, 20, 24)]


def _case(oracle_lib, Nxx, seed):
    import torch
    from nrpy_amd.runtime import Patch, PatchConfig

    P = Patch(PatchConfig(coord=2, Nxx=Nxx, fd_order=6, symmetry_axes=0, enable_KO=True, grid_physical_size=64.0,
                          SINHW=0.2, eta=2.0))
    O = oracle_lib.OracleGrid(2, Nxx, 6, 64.0, 0.2, variant=1)
    y = ID.smooth_perturbation(ID.brill_lindquist(2, O.xx, 64.0, 0.2, (0.6, 0.4), (0.3, 0.2, 0.5),
                                                  (-0.3, -0.2, -0.5)), O.xx, seed)
    O.apply_bcs_outerextrap_and_inner(y)
    y = np.ascontiguousarray(y)
    yd = torch.from_numpy(y).cuda()
    aux = P.new_gfs(6)
    P.Ricci_eval(yd, aux)
    return P, O, y, yd, aux


def test_noalloc_diagnostics_match_the_allocating_entry_point(hip, oracle_lib):
    import torch
    from nrpy_amd.runtime import DIAG_INTEGRANDS, DIAG_RECIPES, _check

    lib = hip.load_library()
    for k, Nxx in enumerate(GRIDS):
        P, O, y, yd, aux = _case(oracle_lib, Nxx, seed=60 + k)
        diag_a, diag_b = P.new_gfs(6), P.new_gfs(6)
        ref = P.diagnostics_constraints(aux, yd, diag_a)
        sums = torch.full((8,), float("nan"), dtype=torch.float64, device="cuda")
        _check(lib.bhah_hip_diagnostics_constraints_noalloc(P.handle, ctypes.c_void_p(aux.data_ptr()),
                                                             ctypes.c_void_p(yd.data_ptr()),
                                                             ctypes.c_void_p(diag_b.data_ptr()),
                                                             ctypes.c_void_p(sums.data_ptr())))
        torch.cuda.synchronize()
        assert torch.equal(diag_a, diag_b)
        s = sums.cpu().numpy()
        for q, recipe in enumerate(DIAG_RECIPES):
            assert s[4 * q] == ref[recipe]["volume"], (Nxx, recipe)
            for j, f in enumerate(DIAG_INTEGRANDS):
                assert s[4 * q + 1 + j] == ref[recipe][f]["integral"], (Nxx, recipe, f)
        oref = O.volume_integrals(O.constraints(y, O.ricci(y)))
        for q, recipe in enumerate(DIAG_RECIPES):
            v = oref[recipe]["volume"]
            assert abs(s[4 * q] - v) <= 1e-12 * v
            for j, f in enumerate(DIAG_INTEGRANDS):
                I = oref[recipe][f]["integral"]
                assert abs(s[4 * q + 1 + j] - I) <= 1e-12 * abs(I), (Nxx, recipe, f)
        assert s[4] > 0  # the r > 8 recipe is not empty on these grids


def test_noalloc_rejects_null_arguments(hip):
    from nrpy_amd.runtime import Patch, PatchConfig

    lib = hip.load_library()
    P = Patch(PatchConfig(coord=2, Nxx=(8, 4, 4), fd_order=6, symmetry_axes=0, enable_KO=True,
                          grid_physical_size=64.0))
    assert lib.bhah_hip_diagnostics_constraints_noalloc(P.handle, None, None, None, None) == -1
    assert b"NULL" in lib.bhah_hip_last_error()
