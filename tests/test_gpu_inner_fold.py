"""The MoL step with the inner-parity boundary condition of k folded into the RK pass against
the step with the separate copy (BHAH_HIP_SEPARATE_INNER_BC=1 at patch creation): the same
state bit for bit, eagerly and through the cached graph, at the default schedules and tile.

The fold needs every pair of points (2q, 2q+1) to find both sources of its k in one 16-byte-
aligned pair; the first case (fd6, four ghost zones, even rows) meets that, which the test
checks from the patch's inner boundary lists, so the folded pass is the one compared.  The
second case (fd4, three ghost zones) does not, and its step keeps the separate copy."""

import os

import numpy as np
import pytest

from nrpy_amd import initial_data as ID

pytestmark = pytest.mark.gpu

CASES = {
    "sinhspherical_fd6": dict(coord=2, Nxx=(20, 12, 10), fd_order=6, symmetry_axes=0, enable_KO=True,
                              grid_physical_size=10.0, SINHW=0.4),
    "spherical_fd4_sym2": dict(coord=1, Nxx=(24, 8, 4), fd_order=4, symmetry_axes=4, enable_KO=False,
                               grid_physical_size=10.0, SINHW=0.2),
}


def pairs_fit(P) -> bool:
    """Whether both points of every pair take k from one aligned pair (the fold's condition)."""
    dst, src, _ = P.inner_bc()
    n = int(np.prod(P.Nx))
    s = np.arange(n, dtype=np.int64)
    s[dst] = src
    return n % 2 == 0 and bool(np.all((s[0::2] >> 1) == (s[1::2] >> 1)))


@pytest.mark.parametrize("name", list(CASES))
def test_folded_inner_bc_step_equals_the_separate_copy(hip, name):
    import torch
    from nrpy_amd.runtime import MoLState, Patch, PatchConfig

    cfg = PatchConfig(eta=2.0, **CASES[name])
    old = os.environ.pop("BHAH_HIP_SEPARATE_INNER_BC", None)
    try:
        PF = Patch(cfg, tuned=False)
        os.environ["BHAH_HIP_SEPARATE_INNER_BC"] = "1"
        PS = Patch(cfg, tuned=False)
    finally:
        os.environ.pop("BHAH_HIP_SEPARATE_INNER_BC", None)
        if old is not None:
            os.environ["BHAH_HIP_SEPARATE_INNER_BC"] = old
    assert pairs_fit(PF) == (name == "sinhspherical_fd6")
    c = CASES[name]
    xx = PF.xx()
    y = ID.brill_lindquist(c["coord"], xx, c["grid_physical_size"], c["SINHW"], (0.5, 0.4), (0.1, 0.2, 0.5),
                           (-0.1, 0, -0.6))
    y = ID.smooth_perturbation(y, xx, 7, amplitude=2e-2)
    y0 = torch.from_numpy(np.ascontiguousarray(y)).cuda()
    PF.apply_bcs_outerextrap_and_inner(y0)
    for use_graph in (False, True):
        SF, SS = MoLState(PF, y0.clone()), MoLState(PS, y0.clone())
        SF.steps(3, PF.dt, use_graph=use_graph)
        SS.steps(3, PS.dt, use_graph=use_graph)
        torch.cuda.synchronize()
        assert torch.isfinite(SF.y_n).all()
        assert torch.equal(SF.y_n, SS.y_n), (name, use_graph, (SF.y_n - SS.y_n).abs().max().item())
        assert not torch.equal(SF.y_n, y0)
