"""The PyTorch transcription used to check the adjoint gradients reproduces the
oracle's RK4 step (it is then an independent differentiable forward map)."""

import numpy as np
import pytest

from nrpy_amd import initial_data as ID


@pytest.mark.slow
def test_torch_step_matches_oracle(oracle_lib):
    import torch
    from tests.torch_transcription import TorchBSSN

    O = oracle_lib.OracleGrid(1, (24, 6, 2), 4, 10.0, 0.2, variant=0)
    y = ID.brill_lindquist(1, O.xx, 10.0, 0.2, (0.5, 0.45), (0, 0, 0.5), (0, 0, -0.55))
    y = np.ascontiguousarray(ID.smooth_perturbation(y, O.xx, 4, amplitude=1e-2))
    dt = O.cfl_dt(0.5)
    T = TorchBSSN(O, dt)
    y1 = T.step(torch.from_numpy(y)).numpy()
    st = O.new_state(y.copy())
    O.bssn_mol_step(st, O.params(eta=2.0), dt)
    ref = st["y_n"]
    num = np.abs(y1 - ref).reshape(24, -1).max(axis=1)
    den = np.maximum(np.abs(ref).reshape(24, -1).max(axis=1), 1e-300)
    assert (num / np.where(den > 0, den, 1)).max() < 1e-11, num / den
