"""Gradients of a lapse-extraction loss through the BSSN evolution, by reverse mode.

The loss is the mean squared difference of the lapse alpha sampled at fixed points
on a sphere r = r_s over the last W of N RK4 steps, against a target run.  The
forward run is the ordinary MoL evolution (``bhah_hip_MoL_step_forward_in_time``);
the gradient comes from the hand-written adjoint of one RK4 step applied backwards
through the run.  The reverse step reads a tape of the forward step's stages
(``bhah_hip_MoL_step_forward_taped`` / ``bhah_hip_MoL_step_adjoint_taped``):

* when the tapes of all N steps fit ``tape_budget_gib`` of device memory, the
  forward run records them all and the backward pass recomputes nothing;
* otherwise the forward run keeps the state every K steps, and the backward pass
  re-runs each K-step segment from its checkpoint, recording that segment's tapes
  (memory O(N/K) states plus K tapes).

Parameters are the Brill-Lindquist masses and puncture
positions on the z axis, theta = (m1, m2, z1, z2); the initial data depend on
them analytically (conformal factor and lapse psi^-2, all else zero), so the
gradient closes with the exact d y_0 / d theta.

``LapseLossProblem.loss(theta)`` is a differentiable torch function
(``torch.autograd.Function``) of a float64 tensor of 4 parameters.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import initial_data as ID
from .runtime import MoLState, Patch, PatchConfig

ALPHA = ID.EVOL_INDEX["alpha"]
CF = ID.EVOL_INDEX["cf"]


@dataclass
class LapseLossConfig:
    Nxx: Tuple[int, int, int] = (96, 16, 4)
    rmax: float = 10.0
    nsteps: int = 400
    window: int = 50
    checkpoint_every: int = 20
    tape_budget_gib: float = 32.0
    r_sample: float = 4.0
    n_sample: int = 16
    CFL: float = 0.5
    eta: float = 2.0
    target: Tuple[float, float, float, float] = (0.52, 0.48, 0.55, -0.55)


def sample_points(P: Patch, cfg: LapseLossConfig):
    """Flat indices and weights of alpha at (r_s, theta_j = (j + 1/2) pi / n, phi-plane of
    the first interior point): cubic Lagrange interpolation in r, theta_j on the grid
    (cell centres of an n-point theta grid).  Returns (idx[n, 4], w[n, 4])."""
    xx = P.xx()
    NG, Nx = P.NG, P.Nx
    r = xx[0]
    i = int(np.searchsorted(r, cfg.r_sample)) - 1  # r[i] <= r_s < r[i+1]
    st = [i - 1, i, i + 1, i + 2]
    w = []
    for a in st:
        c = 1.0
        for b in st:
            if b != a:
                c *= (cfg.r_sample - r[b]) / (r[a] - r[b])
        w.append(c)
    idx, wts = [], []
    for j in range(cfg.n_sample):
        th = (j + 0.5) * math.pi / cfg.n_sample
        i1 = int(np.argmin(np.abs(xx[1] - th)))
        if abs(xx[1][i1] - th) > 1e-12:
            raise ValueError("sample angles must lie on the theta grid (n_sample = Nxx1)")
        i2 = NG
        idx.append([ALPHA * P.Ntot + a + Nx[0] * (i1 + Nx[1] * i2) for a in st])
        wts.append(w)
    return np.array(idx, np.int64), np.array(wts)


def initial_data(xx, theta) -> Tuple[np.ndarray, np.ndarray]:
    """Brill-Lindquist on all points (ghost zones included) and d y0 / d theta for
    theta = (m1, m2, z1, z2): returns (y0[24, ...], dy[4, 24, ...])."""
    m1, m2, z1, z2 = (float(t) for t in theta)
    X, Y, Z = ID.cartesian_coords(1, xx, 0.0, 0.0)
    r1 = np.sqrt(X * X + Y * Y + (Z - z1) ** 2)
    r2 = np.sqrt(X * X + Y * Y + (Z - z2) ** 2)
    psi = 1.0 + m1 / (2 * r1) + m2 / (2 * r2)
    shape = (24,) + psi.shape
    y0 = np.zeros(shape)
    y0[ALPHA] = y0[CF] = psi ** -2
    dpsi = [1 / (2 * r1), 1 / (2 * r2), m1 / 2 * (Z - z1) / r1 ** 3, m2 / 2 * (Z - z2) / r2 ** 3]
    dy = np.zeros((4,) + shape)
    for k in range(4):
        dy[k, ALPHA] = dy[k, CF] = -2 * psi ** -3 * dpsi[k]
    return y0, dy


class LapseLossProblem:
    """Forward evolution, loss and its reverse-mode gradient on one patch."""

    def __init__(self, cfg: LapseLossConfig = LapseLossConfig()):
        import torch

        self.cfg = cfg
        pc = PatchConfig(Nxx=cfg.Nxx, grid_physical_size=cfg.rmax, eta=cfg.eta, CFL_FACTOR=cfg.CFL)
        self.P = Patch(pc)
        self.dt = self.P.dt
        self.xx = self.P.xx()
        idx, w = sample_points(self.P, cfg)
        self.s_idx = torch.from_numpy(idx).cuda()
        self.s_w = torch.from_numpy(w).cuda()
        self.target_samples = self._forward(torch.tensor(cfg.target, dtype=torch.float64), record=True)[1]

    def samples(self, y):
        return (y.reshape(-1)[self.s_idx] * self.s_w).sum(dim=1)

    def _y0(self, theta):
        import torch

        y0, dy = initial_data(self.xx, theta.detach().cpu().numpy())
        return torch.from_numpy(y0).cuda(), dy

    def _forward(self, theta, record=False, checkpoints: Optional[List] = None, tape=None):
        """Run nsteps; returns (loss or None, window samples [W, n]).  With ``tape`` (a
        [nsteps, tape_doubles] device tensor) every step's stages are recorded; with
        ``checkpoints`` the state is kept every ``checkpoint_every`` steps."""
        import torch

        cfg = self.cfg
        y0, _ = self._y0(theta)
        S = MoLState(self.P, y0.clone())
        first = cfg.nsteps - cfg.window + 1
        rec = []
        for n in range(1, cfg.nsteps + 1):
            if checkpoints is not None and (n - 1) % cfg.checkpoint_every == 0:
                checkpoints.append(S.y_n.clone())
            if tape is not None:
                self.P.MoL_step_forward_taped(S.y_n, self.dt, tape[n - 1])
            else:
                S.steps(1, self.dt, use_graph=True)
            if n >= first:
                rec.append(self.samples(S.y_n))
        s = torch.stack(rec)
        if record:
            return None, s
        d = s - self.target_samples
        return (d * d).mean(), s

    def loss_value(self, theta) -> float:
        import torch

        return float(self._forward(torch.as_tensor(theta, dtype=torch.float64))[0])

    def full_tape_fits(self) -> bool:
        return self.cfg.nsteps * self.P.tape_doubles * 8 <= self.cfg.tape_budget_gib * 2**30

    def forward_for_gradient(self, theta):
        """The forward run keeping what the backward pass needs: (loss, window samples, store),
        store being ("tape", tapes) or ("checkpoints", states)."""
        import torch

        if self.full_tape_fits():
            try:
                tape = torch.empty((self.cfg.nsteps, self.P.tape_doubles), dtype=torch.float64, device="cuda")
            except torch.cuda.OutOfMemoryError:
                tape = None  # the device is shared or smaller than the budget: checkpoints instead
            if tape is not None:
                L, s = self._forward(theta, tape=tape)
                return L, s, ("tape", tape)
        ck: List = []
        L, s = self._forward(theta, checkpoints=ck)
        return L, s, ("checkpoints", ck)

    def gradient(self, theta, store, window_samples):
        """dL/dtheta by the adjoint recursion over the taped (or checkpointed) run."""
        import torch

        cfg = self.cfg
        kind, data = store
        first = cfg.nsteps - cfg.window + 1
        dLds = 2.0 * (window_samples - self.target_samples) / window_samples.numel()  # [W, n]
        lam = torch.zeros((24,) + tuple(reversed(self.P.Nx)), dtype=torch.float64, device="cuda")
        flat = lam.view(-1)

        def reverse(n, tape_n):
            if n >= first:
                g = (dLds[n - first][:, None] * self.s_w).reshape(-1)
                flat.index_add_(0, self.s_idx.reshape(-1), g)
            self.P.MoL_step_adjoint_taped(tape_n, self.dt, lam)

        if kind == "tape":
            for n in range(cfg.nsteps, 0, -1):
                reverse(n, data[n - 1])
        else:
            K = cfg.checkpoint_every
            seg = torch.empty((K, self.P.tape_doubles), dtype=torch.float64, device="cuda")
            y = torch.empty_like(data[0])
            for c in range(len(data) - 1, -1, -1):
                n0 = c * K
                n1 = min(n0 + K, cfg.nsteps)
                # re-run steps n0+1 .. n1 from the checkpoint, recording their tapes
                y.copy_(data[c])
                for j in range(n1 - n0):
                    self.P.MoL_step_forward_taped(y, self.dt, seg[j])
                for n in range(n1, n0, -1):
                    reverse(n, seg[n - n0 - 1])
        _, dy = self._y0(theta)
        lam_np = lam.cpu().numpy()
        return np.array([np.sum(lam_np * dy[k]) for k in range(4)])

    def loss(self, theta):
        return _LapseLoss.apply(theta, self)


def _lapse_loss_autograd():
    import torch

    class LapseLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, problem):
            L, s, store = problem.forward_for_gradient(theta.detach())
            ctx.problem, ctx.store, ctx.s = problem, store, s
            ctx.save_for_backward(theta)
            return L.detach().cpu()

        @staticmethod
        def backward(ctx, grad):
            (theta,) = ctx.saved_tensors
            g = ctx.problem.gradient(theta, ctx.store, ctx.s)
            ctx.store = None  # the tapes are released with the graph
            return grad * torch.from_numpy(g).to(theta.dtype), None

    return LapseLoss


class _LazyFn:
    _fn = None

    def apply(self, *a):
        if _LazyFn._fn is None:
            _LazyFn._fn = _lapse_loss_autograd()
        return _LazyFn._fn.apply(*a)


_LapseLoss = _LazyFn()
