"""BSSN evolution throughput on one MI355X: gridpoint-RK4-steps/s, FP64.

Workload (default): full-3D BSSN on a SinhSpherical 192x96x96 grid (AMPL=64,
SINHW=0.2), 6th-order finite differences with Kreiss-Oliger dissipation 0.3,
1+log lapse and Gamma-driver shift (eta=2), RK4 at CFL 0.5, Brill-Lindquist
data with masses 0.6 and 0.4 at (0.3,0.2,0.5) and (-0.3,-0.2,-0.5).

value = interior points x timed RK4 steps / seconds of MoL_step_forward_in_time,
with all fields resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 100] [--warmup 5] [--config c3|c2] [--no-cpu]
                    [--no-batched] [--no-adjoint] [--no-small]
    python bench.py --copy-calibration 512   (plain 512 MiB device copy x3, for
                                              scaling the FETCH_SIZE/WRITE_SIZE counters)

The line also carries "batched": the 64-member ensemble (seeded masses, separations and
grids) through the batched path against the same members stepped one after another
with plain per-patch launches, "adjoint": the lapse-loss gradient (checkpointed
forward plus reverse pass) against the plain forward run, and "wave_c1" / "bssn_c2":
the scalar wave on 64^3 and the BSSN head-on collision on 72x12x2.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    "c3": dict(workload="bssn_sinhspherical_192x96x96_fd6_ko0.3_rk4", coord=2, Nxx=(192, 96, 96), fd_order=6,
               sym=0, KO=True, size=64.0, sinhw=0.2, masses=(0.6, 0.4), pos1=(0.3, 0.2, 0.5),
               pos2=(-0.3, -0.2, -0.5), variant=1),
    "c2": dict(workload="bssn_spherical_72x12x2_fd4_rk4", coord=1, Nxx=(72, 12, 2), fd_order=4, sym=4, KO=False,
               size=7.5, sinhw=0.2, masses=(0.5, 0.5), pos1=(0, 0, 0.5), pos2=(0, 0, -0.5), variant=0),
}
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector (= FP64 matrix) peak, dense
PEAK_HBM_GBS = 8000.0

# Memory-side bytes per rhs_eval launch (all kernels of the schedule), from
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of this bench, scaled by
# the same counters on a plain 512 MiB device copy (FETCH_SIZE x2.000, WRITE_SIZE
# x1.000); summary and commands in profiles/bssn_sinhspherical_192x96x96_hbm_traffic.md.
MEASURED_TRAFFIC = {
    ("SinhSpherical_symnone_KO1_fd6", (192, 96, 96), "p1_lb1_sj_L16x4x4x1_d4_dma_rb1_io_pg_ub_sa_lr40_di_mi_xcd_blk"): {
        "fetch_bytes": 2 * 650252 * 1024.0,
        "write_bytes": 346663 * 1024.0,
    },
    ("SinhSpherical_symnone_KO1_fd6", (192, 96, 96), "p1_lb1_sj_L16x4x4x1_d3_dma_rb1_io_pg_ub_mi_xcd"): {
        "fetch_bytes": 2 * 1054470 * 1024.0,
        "write_bytes": 373350 * 1024.0,
    },
}

# Batched ensemble: 64 independent head-on Brill-Lindquist evolutions on Spherical grids drawn
# uniformly from four sizes, masses m1, m2 ~ U[0.3, 0.7], separation d ~ U[0.6, 2.0] along z
# (punctures at z = +-d/2), rmax = 10, 4th-order FD, RK4, each member at its own CFL timestep.
C4_GRIDS = [(48, 8, 2), (72, 12, 2), (96, 16, 4), (128, 24, 8)]
C4_SEED = 20261202


def c4_members(n: int = 64, seed: int = C4_SEED):
    import numpy as np

    rng = np.random.default_rng(seed)
    m1 = rng.uniform(0.3, 0.7, n)
    m2 = rng.uniform(0.3, 0.7, n)
    d = rng.uniform(0.6, 2.0, n)
    g = rng.integers(0, len(C4_GRIDS), n)
    return [dict(Nxx=C4_GRIDS[g[i]], m1=float(m1[i]), m2=float(m2[i]), d=float(d[i])) for i in range(n)]


def bench_batched(steps: int, warmup: int, n: int = 64) -> dict:
    """Ensemble throughput (sum over members of interior points x steps / s) through the
    batched path, against the same members advanced one after another with plain
    per-patch launches; both end states are compared member by member."""
    import numpy as np
    import torch

    from nrpy_amd import autotune
    from nrpy_amd import initial_data as ID
    from nrpy_amd.runtime import Ensemble, MoLState, Patch, PatchConfig

    mem = c4_members(n)
    cfgs = [PatchConfig(Nxx=m["Nxx"], grid_physical_size=10.0, eta=2.0) for m in mem]
    E = Ensemble(cfgs)
    tuned = autotune.apply_tuned_ensemble(E)
    y0 = []
    for P, st, m in zip(E.members, E.states, mem):
        y = ID.brill_lindquist(1, P.xx(), 10.0, 0.2, (m["m1"], m["m2"]), (0, 0, m["d"] / 2), (0, 0, -m["d"] / 2))
        yd = torch.from_numpy(np.ascontiguousarray(y)).cuda()
        P.apply_bcs_outerextrap_and_inner(yd)
        y0.append(yd)
        st.y_n.copy_(yd)
    seq = []
    for cfg, yd in zip(cfgs, y0):
        Q = Patch(cfg, tuned=False)
        Q.select_schedule(0, E.schedules(0)[E.choice[0]])
        Q.select_schedule(1, E.schedules(1)[E.choice[1]])
        Q.set_tile(*E.tile)
        seq.append(MoLState(Q, yd.clone()))
    npts = sum(c.Nxx[0] * c.Nxx[1] * c.Nxx[2] for c in cfgs)
    stream = torch.cuda.current_stream()

    def timed(fn):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return max(time.perf_counter() - t0, e0.elapsed_time(e1) / 1e3)

    if warmup:
        E.steps(warmup)
        for st, dt in zip(seq, E.dts):
            st.steps(warmup, dt, use_graph=False)
    t_batched = timed(lambda: E.steps(steps))

    def sequential():
        for _ in range(steps):
            for st, dt in zip(seq, E.dts):
                st.step(dt)

    t_seq = timed(sequential)
    err = 0.0
    for st, sq in zip(E.states, seq):
        a, b = st.y_n.cpu().numpy(), sq.y_n.cpu().numpy()
        den = np.abs(b).reshape(24, -1).max(axis=1)
        err = max(err, float((np.abs(a - b).reshape(24, -1).max(axis=1) / np.where(den > 0, den, 1)).max()))
    return {
        "workload": f"bssn_ensemble_{n}_spherical_fd4_rk4 (seed {C4_SEED})",
        "members": n,
        "interior_points": npts,
        "steps": steps,
        "value": npts * steps / t_batched,
        "unit": "gridpoint-RK4-steps/s",
        "sequential_plain_value": npts * steps / t_seq,
        "vs_sequential_plain": t_seq / t_batched,
        "max_rel_diff_vs_sequential": err,
        "schedule": {"rhs": E.schedules(0)[E.choice[0]], "ricci": E.schedules(1)[E.choice[1]],
                     "tile": list(E.tile), "max_workgroups": E.max_wg, "tuned": tuned is not None},
    }


def bench_setup() -> dict:
    """Patch set-up (reference-metric arrays, boundary structs, CFL timestep) on the host loops
    against the device kernels, for the 192x96x96 grid and for the 64 ensemble members."""
    import dataclasses

    import torch

    from nrpy_amd.runtime import Patch, PatchConfig

    c = CONFIGS["c3"]
    c3 = PatchConfig(coord=c["coord"], Nxx=c["Nxx"], fd_order=c["fd_order"], symmetry_axes=c["sym"],
                     enable_KO=c["KO"], grid_physical_size=c["size"], SINHW=c["sinhw"])
    c4 = [PatchConfig(Nxx=m["Nxx"], grid_physical_size=10.0, eta=2.0) for m in c4_members()]
    out = {}
    for name, cfgs in (("bssn_sinhspherical_192x96x96", [c3]), ("ensemble_64_members", c4)):
        for where in (False, True):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ps = [Patch(dataclasses.replace(cf, setup_on_device=where), tuned=False) for cf in cfgs]
            torch.cuda.synchronize()
            out[f"{name}_{'device' if where else 'host'}_s"] = time.perf_counter() - t0
            del ps
    # the ensemble up to its first step: set-up and Brill-Lindquist data of all 64 members, on
    # the host (host loops, numpy data, one copy per member) or on the device (set-up kernels,
    # the closed-form data kernel)
    import numpy as np

    from nrpy_amd import initial_data as ID
    from nrpy_amd.runtime import Ensemble

    mem = c4_members()
    states = {}
    for where in (False, True):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E = Ensemble([dataclasses.replace(cf, setup_on_device=where) for cf in c4])
        for P, st, m in zip(E.members, E.states, mem):
            pos1, pos2 = (0.0, 0.0, m["d"] / 2), (0.0, 0.0, -m["d"] / 2)
            if where:
                P.initial_data_brill_lindquist(st.y_n, m["m1"], pos1, m["m2"], pos2)
            else:
                y = ID.brill_lindquist(1, P.xx(), 10.0, 0.2, (m["m1"], m["m2"]), pos1, pos2)
                st.y_n.copy_(torch.from_numpy(np.ascontiguousarray(y)))
            P.apply_bcs_outerextrap_and_inner(st.y_n)
        torch.cuda.synchronize()
        out[f"ensemble_64_members_to_first_step_{'device' if where else 'host'}_s"] = time.perf_counter() - t0
        states[where] = [st.y_n.clone() for st in E.states]
        E.close()
    out["ensemble_initial_data_max_abs_diff"] = max(float((a - b).abs().max()) for a, b in zip(states[False],
                                                                                               states[True]))
    return out


def _timed(fn):
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def bench_wave(steps: int = 200) -> dict:
    """Scalar wave, Cartesian 64^3, 4th-order FD, RK4 at CFL 0.5 on [-10,10]^3 (200 steps, as the
    reference example); uu = 2 + Gaussian pulse, vv = 0; max error against the exact
    solution inside r < 5 at the end (by then boundary reflections dominate it)."""
    import numpy as np
    import torch

    from nrpy_amd import initial_data as ID
    from nrpy_amd.runtime import CARTESIAN, SYSTEM_WAVE, MoLState, Patch, PatchConfig

    P = Patch(PatchConfig(system=SYSTEM_WAVE, coord=CARTESIAN, Nxx=(64, 64, 64), fd_order=4, symmetry_axes=0,
                          grid_physical_size=10.0, wavespeed=1.0))
    xx = P.xx()
    y0 = ID.wave_spherical_gaussian(xx, 0.0, sigma=1.0)
    S = MoLState(P, torch.from_numpy(np.ascontiguousarray(y0)).cuda())
    S.steps(2, P.dt)
    S.y_n.copy_(torch.from_numpy(np.ascontiguousarray(y0)).cuda())
    S.reset_clock()
    sec = _timed(lambda: S.steps(steps, P.dt))
    exact = ID.wave_spherical_gaussian(xx, steps * P.dt, sigma=1.0)
    X, Y, Z = ID.cartesian_coords(0, xx, 0.0, 0.0)
    mask = np.sqrt(X * X + Y * Y + Z * Z) < 5.0
    err = float(np.abs(S.y_n.cpu().numpy()[0] - exact[0])[mask].max())
    npts = 64 ** 3
    return {"workload": "wave_cartesian_64^3_fd4_rk4_200steps", "value": npts * steps / sec,
            "unit": "gridpoint-RK4-steps/s", "ms_per_step": sec * 1e3 / steps,
            "hbm_GBs_at_272B_per_point_step": 272.0 * npts * steps / sec / 1e9,
            "max_abs_err_vs_exact_r_lt_5": err,
            "note": "t = 31.25: extrapolation-boundary reflections have crossed the whole box"}


def bench_c2(steps: int = 200) -> dict:
    """BSSN head-on collision on Spherical 72x12x2 (C2), hipGraph steps: a 1,728-point grid,
    bound by kernel launch latency rather than arithmetic or bandwidth."""
    import numpy as np
    import torch

    from nrpy_amd import initial_data as ID
    from nrpy_amd.runtime import MoLState, Patch, PatchConfig

    c = CONFIGS["c2"]
    P = Patch(PatchConfig(coord=c["coord"], Nxx=c["Nxx"], fd_order=4, symmetry_axes=4, grid_physical_size=c["size"],
                          eta=2.0))
    y = ID.brill_lindquist(1, P.xx(), c["size"], 0.2, c["masses"], c["pos1"], c["pos2"])
    S = MoLState(P, torch.from_numpy(np.ascontiguousarray(y)).cuda())
    P.apply_bcs_outerextrap_and_inner(S.y_n)
    S.steps(2, P.dt)
    sec = _timed(lambda: S.steps(steps, P.dt))
    npts = c["Nxx"][0] * c["Nxx"][1] * c["Nxx"][2]
    return {"workload": c["workload"], "value": npts * steps / sec, "unit": "gridpoint-RK4-steps/s",
            "us_per_step": sec * 1e6 / steps}


def bench_adjoint() -> dict:
    """Lapse-extraction loss on Spherical 96x16x4 (400 RK4 steps, loss over the last 50,
    16 samples on r = 4, checkpoints every 20): time of the plain forward run against
    the gradient (checkpointed forward + reverse pass with recomputation)."""
    import torch

    from nrpy_amd.adjoint import LapseLossProblem

    prob = LapseLossProblem()
    theta0 = [0.5, 0.5, 0.5, -0.5]
    prob.loss_value(theta0)  # warm-up: graphs, workspaces and the tape's device memory
    warm = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    prob.loss(warm).backward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L = prob.loss_value(theta0)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = prob.loss(theta)
    loss.backward()
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    cfg = prob.cfg
    npts = cfg.Nxx[0] * cfg.Nxx[1] * cfg.Nxx[2]
    return {
        "workload": "bssn_lapse_loss_gradient_spherical_96x16x4_400steps_K20",
        "backward_mode": "all steps taped" if prob.full_tape_fits() else "K-step segments re-run and taped",
        "tape_GB": cfg.nsteps * prob.P.tape_doubles * 8 / 1e9,
        "loss": L,
        "gradient_m1_m2_z1_z2": [float(x) for x in theta.grad],
        "forward_s": t_fwd,
        "gradient_s": t_grad,
        "gradient_vs_forward_time": t_grad / t_fwd,
        "plain_value": npts * cfg.nsteps / t_fwd,
        "adjoint_value": npts * cfg.nsteps / max(t_grad - t_fwd, 1e-9),
        "unit": "gridpoint-RK4-steps/s",
    }


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            return next((line.split(":", 1)[1].strip() for line in f if line.startswith("model name")), "unknown")
    except OSError:
        return "unknown"


def cpu_baseline(c: dict, min_seconds: float = 10.0, max_steps: int = 24) -> dict:
    """The oracle (plain-C restatement) on the host cores: RK4 steps of the same grid until
    about ``min_seconds`` of CPU work (at most ``max_steps``)."""
    import numpy as np

    from nrpy_amd import initial_data as ID
    from oracle import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", min(16, os.cpu_count() or 1)))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    # the restatement built with the reference's project flags (-O2, AVX-512, OpenMP), with
    # rhs_eval and Ricci_eval unrolled and vectorised along xx0 as the reference's generated
    # code is; the slot-table checker build where the host has no AVX-512
    native = oracle.native_supported() and os.path.exists(oracle.NATIVE_LIB_PATH)
    O = oracle.OracleGrid(c["coord"], c["Nxx"], c["fd_order"], c["size"], c["sinhw"], variant=c["variant"],
                          native=native)
    native = native and bool(O.L.og_unrolled_available(c["variant"], c["fd_order"]))
    form = ("unrolled-stencil C, -O2 -march=x86-64-v4 -fopenmp, xx0 vectorised" if native else
            "slot-table C, -O2 -march=x86-64-v3 -fopenmp")
    y = ID.brill_lindquist(c["coord"], O.xx, c["size"], c["sinhw"], c["masses"], c["pos1"], c["pos2"])
    O.apply_bcs_outerextrap_and_inner(y)
    st = O.new_state(np.ascontiguousarray(y))
    dt = O.cfl_dt(0.5)
    t0 = time.perf_counter()
    steps = 0
    while steps < max_steps and (steps == 0 or time.perf_counter() - t0 < min_seconds):
        O.bssn_mol_step(st, O.params(eta=2.0), dt)
        steps += 1
    sec = time.perf_counter() - t0
    npts = c["Nxx"][0] * c["Nxx"][1] * c["Nxx"][2]
    return {
        "value": npts * steps / sec,
        "unit": "gridpoint-RK4-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} RK4 step(s) of the full {c['Nxx'][0]}x{c['Nxx'][1]}x{c['Nxx'][2]} grid, "
                  f"C restatement ({form}; OpenMP, {threads} threads), {sec:.1f} s",
        "host_cpu": _cpu_model(),
    }


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--repeats", type=int, default=5, help="further timed blocks of K steps for the spread")
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-ensemble comparison")
    ap.add_argument("--batched-steps", type=int, default=200)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient timing")
    ap.add_argument("--no-small", action="store_true", help="skip the wave 64^3 and BSSN 72x12x2 lines")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB three times, then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the single-patch path runs on one device; multi-GPU sharding is design-only).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from nrpy_amd import initial_data as ID
    from nrpy_amd.runtime import MoLState, Patch, PatchConfig

    torch.cuda.set_device(0)
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": n * 8, "copies": 3}))
        return
    c = CONFIGS[args.config]
    cfg = PatchConfig(coord=c["coord"], Nxx=c["Nxx"], fd_order=c["fd_order"], symmetry_axes=c["sym"],
                      enable_KO=c["KO"], grid_physical_size=c["size"], SINHW=c["sinhw"], eta=2.0,
                      KO_strength_gauge=0.3, KO_strength_nongauge=0.3)
    P = Patch(cfg)
    xx = P.xx()
    y = ID.brill_lindquist(c["coord"], xx, c["size"], c["sinhw"], c["masses"], c["pos1"], c["pos2"])
    S = MoLState(P, torch.from_numpy(np.ascontiguousarray(y)).cuda())
    P.apply_bcs_outerextrap_and_inner(S.y_n)
    dt = P.dt
    use_graph = not args.no_graph
    if args.warmup:
        S.steps(args.warmup, dt, use_graph=use_graph)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    S.steps(args.steps, dt, use_graph=use_graph)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    sec = max(wall, gpu_ms / 1e3)
    npts = c["Nxx"][0] * c["Nxx"][1] * c["Nxx"][2]
    value = npts * args.steps / sec
    assert torch.isfinite(S.y_n).all().item(), "evolution produced non-finite values"
    # run-to-run spread: further blocks of the same K steps, reported beside `value`
    rep_ms = []
    for _ in range(args.repeats):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        S.steps(args.steps, dt, use_graph=use_graph)
        torch.cuda.synchronize()
        rep_ms.append((time.perf_counter() - t1) * 1e3 / args.steps)

    # Dominant kernel (rhs_eval) timed on its own stream with HIP events.
    # FP64 operations per interior point of the schedules that run, counted on their emitted
    # CSE DAG (add/sub/mul/div 1, fma 2); the tuned schedules are single kernels, so nothing
    # in these counts is recomputed work
    from nrpy_amd.runtime import KERNEL_RHS, KERNEL_RICCI

    with open(os.path.join(ROOT, "nrpy_amd", "csrc", "gen", "kernel_meta.json")) as f:
        cand_flops = json.load(f)["candidates"]
    tag = {1: "Spherical_sym2_KO0", 2: "SinhSpherical_symnone_KO1"}[c["coord"]]
    flops_rhs = cand_flops[f"rhs_{tag}_fd{c['fd_order']}_{P.schedule(KERNEL_RHS)}"]
    flops_ricci = cand_flops[f"ricci_{tag}_fd{c['fd_order']}_{P.schedule(KERNEL_RICCI)}"]
    # the one-thread-per-point schedule's count: every upwind and KO sum formed on its own, as
    # the reference's generated code does
    flops_rhs_expanded = cand_flops[f"rhs_{tag}_fd{c['fd_order']}_p1_lb1"]
    reps = 10
    P.Ricci_eval(S.y_n, S.aux)
    P.rhs_eval(S.aux, S.y_n, S.k_odd)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        P.rhs_eval(S.aux, S.y_n, S.k_odd)
    e1.record(stream)
    torch.cuda.synchronize()
    rhs_ms = e0.elapsed_time(e1) / reps
    e0.record(stream)
    for _ in range(reps):
        P.Ricci_eval(S.y_n, S.aux)
    e1.record(stream)
    torch.cuda.synchronize()
    ricci_ms = e0.elapsed_time(e1) / reps
    achieved = flops_rhs * npts / (rhs_ms * 1e-3) / 1e12
    from nrpy_amd.autotune import build_tag

    sched = P.schedule(KERNEL_RHS)
    measured = MEASURED_TRAFFIC.get((build_tag(cfg), tuple(c["Nxx"]), sched))
    # algorithmic bytes: every input value read once (24 evolved + 6 Ricci gridfunctions,
    # ghost zones included), every right-hand side written once
    ntot = 1
    for n in c["Nxx"]:
        ntot *= n + 2 * (c["fd_order"] // 2 + 1)
    alg_bytes = 8.0 * (30 * ntot + 24 * npts)
    result = {
        "metric": "BSSN gridpoint-RK4-steps/s FP64 on 1 GPU",
        "value": value,
        "unit": "gridpoint-RK4-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Brill-Lindquist initial data computed in-process)",
        "config": {
            "workload": c["workload"],
            "grid": list(c["Nxx"]),
            "interior_points": npts,
            "fd_order": c["fd_order"],
            "dt": dt,
            "hipgraph": use_graph,
        },
        "repeats": {
            "blocks": len(rep_ms),
            "steps_each": args.steps,
            "median_ms_per_step": float(np.median(rep_ms)) if rep_ms else None,
            "min_ms_per_step": min(rep_ms) if rep_ms else None,
            "max_ms_per_step": max(rep_ms) if rep_ms else None,
            "median_value": npts / (float(np.median(rep_ms)) * 1e-3) if rep_ms else None,
        },
        "roofline": {
            # compute roofline: the stencil arithmetic runs on the FP64 vector ALUs (no matrix
            # instructions); MI355X's dense FP64 peak is the same for vector and matrix units
            "bound": "mfma",
            "compute_unit": "FP64 VALU",
            "compute": "FP64 VALU (stencil arithmetic, no MFMA); peak is MI355X dense FP64",
            "kernel": f"rhs_eval: bssn_rhs_{tag}_fd{c['fd_order']}_{sched}_g*",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": (measured["fetch_bytes"] + measured["write_bytes"]) if measured else None,
            "algorithmic_bytes": alg_bytes,
            "flops_per_point": flops_rhs,
            "expanded_flops_per_point": flops_rhs_expanded,
            "expanded_rate_tflops": flops_rhs_expanded * npts / (rhs_ms * 1e-3) / 1e12,
            "avg_launch_ms": rhs_ms,
            "ricci_avg_launch_ms": ricci_ms,
            "ricci_tflops": flops_ricci * npts / (ricci_ms * 1e-3) / 1e12,
        },
    }
    if not args.no_small:
        result["wave_c1"] = bench_wave()
        result["bssn_c2"] = bench_c2()
        result["setup"] = bench_setup()
    if not args.no_batched:
        result["batched"] = bench_batched(args.batched_steps, 2)
    if not args.no_adjoint:
        result["adjoint"] = bench_adjoint()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(c)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
