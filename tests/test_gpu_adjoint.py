"""Reverse mode on the GPU: the adjoint of one RK4 step against central finite
differences of the forward step (dot-product test), and the gradient of the
lapse-extraction loss (96x16x4, 400 steps, window 50, checkpoints every 20)
against central finite differences with h = 1e-6."""

import numpy as np
import pytest

from nrpy_amd import initial_data as ID

pytestmark = pytest.mark.gpu


def test_step_adjoint_dot_product(hip):
    import torch
    from nrpy_amd.runtime import MoLState, Patch, PatchConfig

    P = Patch(PatchConfig(Nxx=(24, 8, 4), grid_physical_size=10.0, eta=2.0), tuned=False)
    xx = P.xx()
    y = ID.brill_lindquist(1, xx, 10.0, 0.2, (0.5, 0.4), (0, 0, 0.5), (0, 0, -0.6))
    y = ID.smooth_perturbation(y, xx, 3, amplitude=2e-2)
    y[ID.EVOL_INDEX["vetU0"]] += 0.05  # keep beta^r away from the upwind switch
    y[ID.EVOL_INDEX["vetU1"]] += 0.05
    yd = torch.from_numpy(np.ascontiguousarray(y)).cuda()
    P.apply_bcs_outerextrap_and_inner(yd)
    rng = np.random.default_rng(11)
    v = torch.from_numpy(rng.standard_normal(y.shape)).cuda()
    lam = torch.from_numpy(rng.standard_normal(y.shape)).cuda()
    h = 1e-6

    def step(z):
        S = MoLState(P, z.clone())
        S.step(P.dt)
        return S.y_n

    fd = (step(yd + h * v) - step(yd - h * v)) / (2 * h)
    lhs = float((lam * fd).sum())
    la = lam.clone()
    P.MoL_step_adjoint(yd, P.dt, la)
    rhs = float((la * v).sum())
    print("step adjoint dot product", lhs, rhs)
    assert abs(lhs - rhs) < 1e-6 * abs(lhs), (lhs, rhs)


def test_taped_step_equals_the_mol_step_and_the_recomputing_adjoint(hip):
    """The taped forward step gives the MoL step's state bit for bit, and the adjoint read
    from its tape equals the adjoint that recomputes the stages, bit for bit (same kernels,
    same inputs)."""
    import torch
    from nrpy_amd.runtime import MoLState, Patch, PatchConfig

    P = Patch(PatchConfig(Nxx=(24, 8, 4), grid_physical_size=10.0, eta=2.0), tuned=False)
    xx = P.xx()
    y = ID.brill_lindquist(1, xx, 10.0, 0.2, (0.5, 0.4), (0, 0, 0.5), (0, 0, -0.6))
    y = ID.smooth_perturbation(y, xx, 5, amplitude=2e-2)
    yd = torch.from_numpy(np.ascontiguousarray(y)).cuda()
    P.apply_bcs_outerextrap_and_inner(yd)
    S = MoLState(P, yd.clone())
    S.step(P.dt)
    tape = torch.full((P.tape_doubles,), float("nan"), dtype=torch.float64, device="cuda")
    yt = yd.clone()
    P.MoL_step_forward_taped(yt, P.dt, tape)
    torch.cuda.synchronize()
    assert torch.equal(yt, S.y_n)
    nv = yd.numel()
    # states and right-hand sides at every point; Ricci_eval writes RbarDD at interior points
    assert torch.isfinite(tape[: 8 * nv]).all()
    NG = P.NG
    aux = tape[8 * nv:].view(4 * 6, *yd.shape[1:])
    assert torch.isfinite(aux[:, NG:-NG, NG:-NG, NG:-NG]).all()
    assert torch.equal(tape[:nv].view_as(yd), yd)  # the tape starts with y_n
    rng = np.random.default_rng(5)
    lam = torch.from_numpy(rng.standard_normal(y.shape)).cuda()
    la, lb = lam.clone(), lam.clone()
    P.MoL_step_adjoint(yd, P.dt, la)
    P.MoL_step_adjoint_taped(tape, P.dt, lb)
    torch.cuda.synchronize()
    assert torch.equal(la, lb)
    with pytest.raises(Exception):
        P.MoL_step_adjoint_taped(tape[:-1], P.dt, lb)


def test_checkpointed_gradient_equals_the_full_tape(hip):
    """Both backward modes (all steps taped in the forward run, or K-step segments re-run from
    checkpoints and taped) give the same gradient bit for bit: the segments re-run the same
    kernels on the same states."""
    import torch
    from nrpy_amd.adjoint import LapseLossConfig, LapseLossProblem

    cfg = LapseLossConfig(Nxx=(24, 6, 2), nsteps=11, window=5, checkpoint_every=4, n_sample=6)
    full = LapseLossProblem(cfg)
    assert full.full_tape_fits()
    seg = LapseLossProblem(LapseLossConfig(**{**cfg.__dict__, "tape_budget_gib": 1e-6}))
    assert not seg.full_tape_fits()
    grads = []
    for prob in (full, seg):
        theta = torch.tensor([0.5, 0.45, 0.5, -0.55], dtype=torch.float64, requires_grad=True)
        prob.loss(theta).backward()
        grads.append(theta.grad.numpy().copy())
    print("full tape", grads[0], "segments", grads[1])
    assert np.array_equal(grads[0], grads[1]), grads


def test_lapse_loss_gradient_matches_central_differences(hip):
    import torch
    from nrpy_amd.adjoint import LapseLossProblem

    prob = LapseLossProblem()
    theta = torch.tensor([0.5, 0.5, 0.5, -0.5], dtype=torch.float64, requires_grad=True)
    L = prob.loss(theta)
    L.backward()
    g = theta.grad.numpy()
    h = 1e-6
    fd = []
    for k in range(4):
        tp = theta.detach().numpy().copy()
        tm = tp.copy()
        tp[k] += h
        tm[k] -= h
        fd.append((prob.loss_value(tp) - prob.loss_value(tm)) / (2 * h))
    fd = np.array(fd)
    print("loss", float(L), "adjoint", g, "central FD", fd)
    assert float(L) > 0
    assert np.all(np.abs(g - fd) <= 1e-6 * np.abs(fd) + 1e-12 * np.abs(fd).max()), (g, fd)


def test_gradient_matches_torch_transcription(hip):
    """On a 24x6x2 patch (8 steps, window 4, checkpoints every 3), the adjoint gradient
    equals torch autograd through the PyTorch transcription of the same step."""
    import torch
    from nrpy_amd.adjoint import LapseLossConfig, LapseLossProblem
    from tests.torch_transcription import TorchBSSN

    cfg = LapseLossConfig(Nxx=(24, 6, 2), nsteps=8, window=4, checkpoint_every=3, n_sample=6)
    prob = LapseLossProblem(cfg)
    theta = torch.tensor([0.5, 0.45, 0.5, -0.55], dtype=torch.float64, requires_grad=True)
    L = prob.loss(theta)
    L.backward()
    g = theta.grad.numpy().copy()

    T = TorchBSSN(prob.P, prob.dt)
    idx, w = prob.s_idx.cpu(), prob.s_w.cpu()
    from nrpy_amd.initial_data import cartesian_coords

    X, Y, Z = (torch.from_numpy(a) for a in cartesian_coords(1, prob.xx, 0.0, 0.0))
    ia, ic = 6, 10  # alpha, cf

    def run(th):
        r1 = torch.sqrt(X * X + Y * Y + (Z - th[2]) ** 2)
        r2 = torch.sqrt(X * X + Y * Y + (Z - th[3]) ** 2)
        psi = 1.0 + th[0] / (2 * r1) + th[1] / (2 * r2)
        rows = [torch.zeros_like(psi)] * 24
        rows[ia] = rows[ic] = psi ** -2
        y = torch.stack(rows)
        rec = []
        for n in range(1, cfg.nsteps + 1):
            y = T.step(y)
            if n >= cfg.nsteps - cfg.window + 1:
                rec.append((y.reshape(-1)[idx] * w).sum(1))
        return torch.stack(rec)

    with torch.no_grad():
        target = run(torch.tensor(cfg.target, dtype=torch.float64))
    th = theta.detach().clone().requires_grad_(True)
    s = run(th)
    Lt = ((s - target) ** 2).mean()
    Lt.backward()
    gt = th.grad.numpy()
    print("loss", float(L), float(Lt), "adjoint", g, "torch autograd", gt)
    assert abs(float(L) - float(Lt)) < 1e-9 * float(Lt)
    assert np.all(np.abs(g - gt) <= 1e-7 * np.abs(gt).max()), (g, gt)
