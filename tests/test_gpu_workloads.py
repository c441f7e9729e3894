"""GPU parity on the benchmark workloads themselves, through the C-ABI, against the oracle.

* The 192x96x96 SinhSpherical fd6 + KO grid with the tuned workgroup-tiled schedules, whose
  16x4x4 tiles all are full there and whose tile count (6,912) is a multiple of 8, so the
  XCD-aware tile order is on; and a 32x16x16 grid (32 full tiles) that runs the same code
  paths in seconds.  Per call at 1e-12, RK4 steps at 1e-10.
* The scalar wave on 64^3 for 200 RK4 steps (against the oracle and the exact solution).
* The 64-member batched ensemble (seeded masses, separations and grids of four sizes) for a
  few steps, every member against its own oracle run.
* The head-on collision on Spherical 72x12x2 to t = 5 with the reference driver's output
  cadence (every 0.25): fields and the constraint norms at each output time.

Each test prints the errors it reached; the tolerances are the north-star bars (1e-12
relative per call, 1e-10 relative after steps), relative to each gridfunction's largest
magnitude.
"""

import math

import numpy as np
import pytest

from nrpy_amd import initial_data as ID

pytestmark = pytest.mark.gpu

C3 = dict(size=64.0, sinhw=0.2, masses=(0.6, 0.4), pos1=(0.3, 0.2, 0.5), pos2=(-0.3, -0.2, -0.5))


def rel_err(a, b):
    num = np.abs(a - b).reshape(a.shape[0], -1).max(axis=1)
    den = np.abs(b).reshape(b.shape[0], -1).max(axis=1)
    return num / np.where(den > 0, den, 1.0)


def interior(a, NG):
    return a[:, NG:-NG, NG:-NG, NG:-NG]


def _c3_pair(oracle_lib, Nxx, seed):
    from nrpy_amd.runtime import Patch, PatchConfig

    cfg = PatchConfig(coord=2, Nxx=Nxx, fd_order=6, symmetry_axes=0, enable_KO=True, grid_physical_size=C3["size"],
                      SINHW=C3["sinhw"], eta=2.0, KO_strength_gauge=0.3, KO_strength_nongauge=0.3)
    P = Patch(cfg)
    O = oracle_lib.OracleGrid(2, Nxx, 6, C3["size"], C3["sinhw"], variant=1)
    y = ID.brill_lindquist(2, O.xx, C3["size"], C3["sinhw"], C3["masses"], C3["pos1"], C3["pos2"])
    y = ID.smooth_perturbation(y, O.xx, seed)
    O.apply_bcs_outerextrap_and_inner(y)
    return P, O, np.ascontiguousarray(y)


def _tiles(Nxx, sched):
    """Tile count and whether every tile is full, for an _L<TX>x<TY>x<TZ> schedule."""
    tok = next(t for t in sched.split("_") if t[:1] == "L" and t[1:2].isdigit())
    T = [int(v) for v in tok[1:].split("x")[:3]]
    return math.prod(-(-n // t) for n, t in zip(Nxx, T)), all(n % t == 0 for n, t in zip(Nxx, T))


@pytest.mark.parametrize("Nxx", [(32, 16, 16), (192, 96, 96)])
def test_sinhspherical_fd6_tuned_kernels_per_call(hip, oracle_lib, Nxx):
    import torch
    from nrpy_amd.runtime import KERNEL_RHS, KERNEL_RICCI

    P, O, y = _c3_pair(oracle_lib, Nxx, seed=31)
    for kind in (KERNEL_RHS, KERNEL_RICCI):
        name = P.schedule(kind)
        n, full = _tiles(Nxx, name)
        # the tuned DMA-staged schedules with the XCD tile order on (and the blocked order
        # where the schedule has it: 2 x 4 x 4 tiles divide both grids' tile counts)
        assert "_dma" in name and "_xcd" in name and n % 8 == 0 and full, (name, n, full)
    print(f"\n{Nxx}: rhs {P.schedule(KERNEL_RHS)}, Ricci {P.schedule(KERNEL_RICCI)}")
    aux_o = O.ricci(y)
    rhs_o = O.rhs(y, aux_o, O.params())
    yd = torch.from_numpy(y).cuda()
    aux_d = P.new_gfs(6)
    P.Ricci_eval(yd, aux_d)
    rhs_d = P.new_gfs()
    P.rhs_eval(torch.from_numpy(aux_o).cuda(), yd, rhs_d)
    torch.cuda.synchronize()
    NG = O.NG
    e_ric = rel_err(interior(aux_d.cpu().numpy(), NG), interior(aux_o, NG))
    e_rhs = rel_err(interior(rhs_d.cpu().numpy(), NG), interior(rhs_o, NG))
    print(f"{Nxx}: Ricci max rel err {e_ric.max():.3e}, rhs max rel err {e_rhs.max():.3e}")
    assert e_ric.max() < 1e-12, e_ric
    assert e_rhs.max() < 1e-12, e_rhs


@pytest.mark.parametrize("Nxx,nsteps", [((32, 16, 16), 4), ((192, 96, 96), 2)])
def test_sinhspherical_fd6_rk4_steps(hip, oracle_lib, Nxx, nsteps):
    import torch
    from nrpy_amd.runtime import MoLState

    P, O, y = _c3_pair(oracle_lib, Nxx, seed=32)
    st_o = O.new_state(y)
    dt = P.dt
    assert abs(dt - O.cfl_dt(0.5)) <= 1e-15 * dt
    for _ in range(nsteps):
        O.bssn_mol_step(st_o, O.params(), dt)
    S = MoLState(P, torch.from_numpy(y.copy()).cuda())
    S.steps(nsteps, dt, use_graph=True)
    torch.cuda.synchronize()
    er = rel_err(S.y_n.cpu().numpy(), st_o["y_n"])
    print(f"\n{Nxx}, {nsteps} RK4 steps: max rel err {er.max():.3e} (gf {int(er.argmax())})")
    assert er.max() < 1e-10, er
    assert S.nn == nsteps and S.time == nsteps * dt


def test_wave_64cubed_200_steps(hip, oracle_lib):
    """The wave workload as benchmarked: 64^3, [-10,10]^3, CFL 0.5, 200 steps.  This problem
    amplifies rounding: a 1e-16 relative perturbation of the oracle's own initial data grows
    to ~1e-11 after 25 steps and ~4e-8 (uu) / ~7e-8 (vv) after 200, through the quadratic-
    extrapolation outer boundary.  So the fields must agree to 1e-10 after 10 steps, and after
    200 steps to within 4x the spread the oracle shows under that perturbation."""
    import torch
    from nrpy_amd.runtime import CARTESIAN, SYSTEM_WAVE, MoLState, Patch, PatchConfig

    N, steps, early = 64, 200, 10
    P = Patch(PatchConfig(system=SYSTEM_WAVE, coord=CARTESIAN, Nxx=(N, N, N), fd_order=4, symmetry_axes=0,
                          grid_physical_size=10.0, wavespeed=1.0))
    O = oracle_lib.OracleGrid(0, (N, N, N), 4, 10.0)
    y = ID.wave_spherical_gaussian(O.xx, 0.0, sigma=1.0)
    dt = P.dt
    assert dt == 0.15625
    rng = np.random.default_rng(0)
    y_pert = y * (1 + 1e-16 * rng.standard_normal(y.shape))
    st, sp_ = O.new_state(y, nvars=2), O.new_state(y_pert, nvars=2)
    S = MoLState(P, torch.from_numpy(y.copy()).cuda())
    for n in range(1, steps + 1):
        O.wave_mol_step(st, dt)
        O.wave_mol_step(sp_, dt)
        S.steps(1, dt, use_graph=True)
        if n == early:
            torch.cuda.synchronize()
            er_early = rel_err(S.y_n.cpu().numpy(), st["y_n"])
    torch.cuda.synchronize()
    got = S.y_n.cpu().numpy()
    er = rel_err(got, st["y_n"])
    spread = rel_err(sp_["y_n"], st["y_n"])
    exact = ID.wave_spherical_gaussian(O.xx, steps * dt, sigma=1.0)
    X, Y, Z = ID.cartesian_coords(0, O.xx, 0.0, 0.0)
    mask = np.sqrt(X * X + Y * Y + Z * Z) < 5.0
    e_gpu = np.abs(got[0] - exact[0])[mask].max()
    e_ora = np.abs(st["y_n"][0] - exact[0])[mask].max()
    print(f"\nwave 64^3: step {early} max rel err vs oracle {er_early.max():.3e}; step {steps}: {er} against the "
          f"oracle's own rounding spread {spread}; |uu - exact| (r<5): GPU {e_gpu:.3e}, oracle {e_ora:.3e}")
    assert er_early.max() < 1e-10, er_early
    assert (er <= 4 * spread + 1e-14).all(), (er, spread)
    assert abs(e_gpu - e_ora) <= 1e-6 * e_ora


def test_ensemble_64_members_against_oracle(hip, oracle_lib):
    """The benchmarked ensemble (64 seeded members over the four grid sizes, tuned batched
    schedules, hipGraph steps), every member against its own oracle evolution."""
    import torch
    from bench import C4_GRIDS, c4_members
    from nrpy_amd import autotune
    from nrpy_amd.runtime import Ensemble, PatchConfig

    nsteps = 3
    mem = c4_members(64)
    assert {m["Nxx"] for m in mem} == set(C4_GRIDS)
    cfgs = [PatchConfig(Nxx=m["Nxx"], grid_physical_size=10.0, eta=2.0) for m in mem]
    E = Ensemble(cfgs)
    assert autotune.apply_tuned_ensemble(E) is not None
    ys = []
    for P, st, m in zip(E.members, E.states, mem):
        y = ID.brill_lindquist(1, P.xx(), 10.0, 0.2, (m["m1"], m["m2"]), (0, 0, m["d"] / 2), (0, 0, -m["d"] / 2))
        y = np.ascontiguousarray(ID.smooth_perturbation(y, P.xx(), 40, amplitude=1e-3))
        yd = torch.from_numpy(y).cuda()
        P.apply_bcs_outerextrap_and_inner(yd)
        st.y_n.copy_(yd)
        ys.append(yd.cpu().numpy())
    E.steps(nsteps, use_graph=True)
    torch.cuda.synchronize()
    worst = 0.0
    for k, (m, st, dt) in enumerate(zip(mem, E.states, E.dts)):
        O = oracle_lib.OracleGrid(1, m["Nxx"], 4, 10.0, 0.2, variant=0)
        assert abs(dt - O.cfl_dt(0.5)) <= 1e-15 * dt
        ost = O.new_state(ys[k])
        for _ in range(nsteps):
            O.bssn_mol_step(ost, O.params(eta=2.0), dt)
        er = rel_err(st.y_n.cpu().numpy(), ost["y_n"]).max()
        worst = max(worst, er)
        assert er < 1e-10, (k, m, er)
    print(f"\n64-member ensemble, {nsteps} steps: worst member max rel err {worst:.3e}")


def test_head_on_collision_to_t5_with_constraint_norms(hip, oracle_lib):
    """Spherical 72x12x2, Brill-Lindquist m = 0.5 at z = +-0.5, eta = 2, RK4 at CFL 0.5 until
    time >= 5 (the reference's main loop condition), with the diagnostics driver called before
    every step at output cadence 0.25 (two_blackholes_collide.py:82): the fields at t = 5, all
    six diagnostic gridfunctions there, and the volume integrals of every output step against
    the oracle."""
    import torch
    from nrpy_amd.runtime import MoLState, Patch, PatchConfig

    Nxx, t_final, every = (72, 12, 2), 5.0, 0.25
    P = Patch(PatchConfig(coord=1, Nxx=Nxx, fd_order=4, symmetry_axes=4, enable_KO=False, grid_physical_size=7.5,
                          SINHW=0.2, eta=2.0))
    O = oracle_lib.OracleGrid(1, Nxx, 4, 7.5, 0.2, variant=0)
    y = ID.brill_lindquist(1, O.xx, 7.5, 0.2, (0.5, 0.5), (0, 0, 0.5), (0, 0, -0.5))
    O.apply_bcs_outerextrap_and_inner(y)
    y = np.ascontiguousarray(y)
    st_o = O.new_state(y)
    S = MoLState(P, torch.from_numpy(y.copy()).cuda())
    dt = P.dt
    aux, diag = P.new_gfs(6), P.new_gfs(6)
    series = {f: ([], []) for f in ("H", "MSQUARED", "LAMBDA_CONSTRAINT")}
    outputs, n = 0, 0
    while S.time < t_final:
        vi = P.diagnostics(S.clock, every, S.y_n, aux, diag)
        t_o = n * dt
        out_o = abs(round(t_o / every) * every - t_o) < 0.5 * dt
        assert (vi is not None) == out_o, (n, S.time)
        if vi is not None:
            outputs += 1
            ref = O.volume_integrals(O.constraints(st_o["y_n"], O.ricci(st_o["y_n"])))
            for f, (a, b) in series.items():
                a.append(vi["whole_domain"][f]["L2"])
                b.append(ref["whole_domain"][f]["L2"])
        S.steps(1, dt, use_graph=True)
        O.bssn_mol_step(st_o, O.params(), dt)
        n += 1
    torch.cuda.synchronize()
    # output times 0, 0.25, ..., 5.0: the last step starts within dt/2 of t = 5
    want = sum(1 for k in range(n) if abs(round(k * dt / every) * every - k * dt) < 0.5 * dt)
    assert n == math.ceil(t_final / dt - 1e-12) and outputs == want == 21, (n, outputs, want)
    er = rel_err(S.y_n.cpu().numpy(), st_o["y_n"])
    aux_o = O.ricci(st_o["y_n"])
    c_o = O.constraints(st_o["y_n"], aux_o)
    P.Ricci_eval(S.y_n, aux)
    P.diagnostics_constraints(aux, S.y_n, diag, integrals=False)
    ec = rel_err(interior(diag.cpu().numpy(), O.NG), interior(c_o, O.NG))
    # each norm series against the oracle's, relative to the series' largest value (at t = 0
    # M^2 and the Lambda constraint of Brill-Lindquist data are zero)
    worst_norm = max(np.abs(np.subtract(a, b)).max() / np.abs(b).max() for a, b in series.values())
    print(f"\nhead-on to t = {S.time:.6f} ({n} steps): fields max rel err {er.max():.3e}; constraints "
          f"H, MU0-2, M^2, Lambda max rel err {np.array2string(ec, precision=2)}; norm series worst rel "
          f"err {worst_norm:.3e} over {outputs} outputs")
    assert er.max() < 1e-10, er
    assert ec.max() < 1e-9, ec
    assert worst_norm < 1e-10
