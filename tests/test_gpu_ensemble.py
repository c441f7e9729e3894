"""Batched ensembles on the GPU: every member of a batched evolution equals the
same patch evolved alone through the single-patch path (which the parity tests
pin to the oracle), and a few steps match the oracle directly."""

import numpy as np
import pytest

from nrpy_amd import initial_data as ID

pytestmark = pytest.mark.gpu

GRIDS = [(48, 8, 2), (24, 6, 2), (32, 8, 4)]


def _cfgs():
    from nrpy_amd.runtime import PatchConfig

    return [PatchConfig(Nxx=N, grid_physical_size=10.0, eta=2.0) for N in GRIDS]


def _initial(P, k):
    y = ID.brill_lindquist(1, P.xx(), 10.0, 0.2, (0.4 + 0.05 * k, 0.5), (0, 0, 0.4 + 0.1 * k), (0, 0, -0.5))
    return np.ascontiguousarray(ID.smooth_perturbation(y, P.xx(), 20 + k, amplitude=1e-3))


def rel_err(a, b):
    num = np.abs(a - b).reshape(a.shape[0], -1).max(axis=1)
    den = np.abs(b).reshape(b.shape[0], -1).max(axis=1)
    return num / np.where(den > 0, den, 1.0)


@pytest.mark.parametrize("use_graph", [True, False])
def test_ensemble_matches_single_patch_and_oracle(hip, oracle_lib, use_graph):
    import torch
    from nrpy_amd.runtime import Ensemble, MoLState, Patch

    E = Ensemble(_cfgs())
    E.set_launch(64, 4, 0)
    singles = []
    for k, (cfg, P, st) in enumerate(zip(E.cfgs, E.members, E.states)):
        y = torch.from_numpy(_initial(P, k)).cuda()
        P.apply_bcs_outerextrap_and_inner(y)
        st.y_n.copy_(y)
        Q = Patch(cfg, tuned=False)
        Q.select_schedule(0, E.schedules(0)[0])
        Q.select_schedule(1, E.schedules(1)[0])
        S = MoLState(Q, y.clone())
        singles.append((Q, S))
    E.steps(3, use_graph=use_graph)
    for (Q, S), dt in zip(singles, E.dts):
        S.steps(3, dt, use_graph=False)
    torch.cuda.synchronize()
    for k, ((Q, S), st) in enumerate(zip(singles, E.states)):
        er = rel_err(st.y_n.cpu().numpy(), S.y_n.cpu().numpy())
        assert er.max() < 1e-13, (k, er)
    # member 1 against the oracle directly
    P = E.members[1]
    O = oracle_lib.OracleGrid(1, GRIDS[1], 4, 10.0, 0.2, variant=0)
    y0 = _initial(P, 1)
    O.apply_bcs_outerextrap_and_inner(y0)
    ost = O.new_state(y0)
    for _ in range(3):
        O.bssn_mol_step(ost, O.params(eta=2.0), E.dts[1])
    er = rel_err(E.states[1].y_n.cpu().numpy(), ost["y_n"])
    assert er.max() < 1e-10, er


def test_every_ensemble_schedule_and_launch_shape(hip):
    import torch
    from nrpy_amd.runtime import KERNEL_RHS, KERNEL_RICCI, Ensemble

    E = Ensemble(_cfgs())
    init = []
    for k, (P, st) in enumerate(zip(E.members, E.states)):
        y = torch.from_numpy(_initial(P, k)).cuda()
        P.apply_bcs_outerextrap_and_inner(y)
        init.append(y)
    ref = None
    for ric in E.schedules(KERNEL_RICCI):
        for rhs in E.schedules(KERNEL_RHS):
            for launch in [(64, 4, 0), (256, 1, 0), (64, 4, 7), (32, 8, 64)]:
                E.select_schedule(KERNEL_RICCI, ric)
                E.select_schedule(KERNEL_RHS, rhs)
                E.set_launch(*launch)
                for st, y in zip(E.states, init):
                    st.y_n.copy_(y)
                E.steps(2)
                torch.cuda.synchronize()
                got = [st.y_n.cpu().numpy() for st in E.states]
                if ref is None:
                    ref = got
                    continue
                for k in range(len(got)):
                    er = rel_err(got[k], ref[k])
                    assert er.max() < 1e-12, (ric, rhs, launch, k, er)
