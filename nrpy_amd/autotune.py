"""On-device autotuner for the BSSN kernel schedules and thread-block tiles.

For one workload (coordinate system, grid, FD order, KO) it times every
generated schedule of ``rhs_eval`` and ``Ricci_eval`` (output partitions x
register budgets) and every interior tile shape with HIP events on the
patch's stream, on Brill-Lindquist data perturbed so all terms are live, and
stores the fastest choice in ``nrpy_amd/autotune/<build>_<N0>x<N1>x<N2>.json``.
``Patch`` applies a stored choice automatically (see ``apply_tuned``).

    python -m nrpy_amd.autotune --config c3
    python -m nrpy_amd.autotune --config c4     (batched ensemble: schedules and launch shape)
"""

from __future__ import annotations

import argparse
import json
import os
from typing import Dict, Optional

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_DIR = os.path.join(HERE, "autotune")
TILES = [(64, 4), (128, 2), (256, 1), (32, 8)]


def build_tag(cfg) -> str:
    """Name of the generated kernel build serving a patch configuration."""
    coord = {1: "Spherical", 2: "SinhSpherical", 0: "Cartesian"}[cfg.coord]
    axes = "".join(str(d) for d in range(3) if cfg.symmetry_axes & (1 << d)) or "none"
    return f"{coord}_sym{axes}_KO{int(bool(cfg.enable_KO))}_fd{cfg.fd_order}"


def table_path(cfg) -> str:
    n = "x".join(str(int(v)) for v in cfg.Nxx)
    return os.path.join(TABLE_DIR, f"{build_tag(cfg)}_{n}.json")


def load_choice(cfg) -> Optional[Dict]:
    """Stored choice for this exact grid, else the one for the nearest grid of the same build."""
    p = table_path(cfg)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    if not os.path.isdir(TABLE_DIR):
        return None
    tag = build_tag(cfg)
    npts = cfg.Nxx[0] * cfg.Nxx[1] * cfg.Nxx[2]
    best, bestd = None, None
    for fn in os.listdir(TABLE_DIR):
        if fn.startswith(tag + "_") and fn.endswith(".json"):
            with open(os.path.join(TABLE_DIR, fn)) as f:
                d = json.load(f)
            n = d["Nxx"][0] * d["Nxx"][1] * d["Nxx"][2]
            dist = abs(n - npts) / max(n, npts)
            if bestd is None or dist < bestd:
                best, bestd = d, dist
    return best


def apply_tuned(patch) -> Optional[Dict]:
    """Select the stored schedules and tile for ``patch`` (no-op without a table)."""
    from .runtime import KERNEL_RHS, KERNEL_RICCI, SYSTEM_BSSN

    if patch.cfg.system != SYSTEM_BSSN:
        return None
    d = load_choice(patch.cfg)
    if not d:
        return None
    from .runtime import BHaHError

    for kind, key in ((KERNEL_RHS, "rhs"), (KERNEL_RICCI, "ricci")):
        if d[key] in patch.schedules(kind):
            try:
                patch.select_schedule(kind, d[key])
            except BHaHError:
                pass  # schedule not usable at this grid size: keep the default
    patch.set_tile(*d["tile"])
    return d


def _time(fn, stream, reps: int) -> float:
    import torch

    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(cfg, reps: int = 5, write: bool = True, log=print) -> Dict:
    import numpy as np
    import torch

    from . import initial_data as ID
    from .runtime import KERNEL_RHS, KERNEL_RICCI, BHaHError, Patch

    P = Patch(cfg)
    xx = P.xx()
    y = ID.brill_lindquist(cfg.coord, xx, cfg.grid_physical_size, cfg.SINHW)
    y = ID.smooth_perturbation(y, xx, seed=1)
    yd = torch.from_numpy(np.ascontiguousarray(y)).cuda()
    aux = P.new_gfs(6)
    rhs = P.new_gfs()
    stream = torch.cuda.current_stream()
    timings: Dict[str, float] = {}
    best = {}
    for kind, label in ((KERNEL_RICCI, "ricci"), (KERNEL_RHS, "rhs")):
        fn = (lambda: P.Ricci_eval(yd, aux)) if kind == KERNEL_RICCI else (lambda: P.rhs_eval(aux, yd, rhs))
        res = {}
        for name in P.schedules(kind):
            try:
                P.select_schedule(kind, name)
            except BHaHError as e:  # schedule not usable on this grid
                log(f"{label} {name}: skipped ({e})")
                continue
            res[name] = _time(fn, stream, reps)
            timings[f"{label}:{name}"] = res[name]
            log(f"{label} {name}: {res[name]:.3f} ms")
        best[label] = min(res, key=res.get)
        P.select_schedule(kind, best[label])
    tile_res = {}
    for t in TILES:
        P.set_tile(*t)
        tile_res[t] = _time(lambda: (P.Ricci_eval(yd, aux), P.rhs_eval(aux, yd, rhs)), stream, reps)
        timings[f"tile:{t[0]}x{t[1]}"] = tile_res[t]
        log(f"tile {t}: {tile_res[t]:.3f} ms (ricci+rhs)")
    tile = min(tile_res, key=tile_res.get)
    out = {
        "build": build_tag(cfg),
        "Nxx": list(int(v) for v in cfg.Nxx),
        "rhs": best["rhs"],
        "ricci": best["ricci"],
        "tile": list(tile),
        "timings_ms": timings,
        "device": torch.cuda.get_device_name(0),
    }
    if write:
        os.makedirs(TABLE_DIR, exist_ok=True)
        with open(table_path(cfg), "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
    P.close()
    return out


ENSEMBLE_LAUNCHES = [(64, 4, 0), (256, 1, 0), (128, 2, 0), (64, 4, 1024), (256, 1, 1024), (64, 4, 512),
                     (256, 1, 512)]


def ensemble_table_path(E) -> str:
    n = len(E.members)
    cells = sum(P.cfg.Nxx[0] * P.cfg.Nxx[1] * P.cfg.Nxx[2] for P in E.members)
    return os.path.join(TABLE_DIR, f"ensemble_{build_tag(E.cfgs[0])}_{n}x{cells}.json")


def apply_tuned_ensemble(E) -> Optional[Dict]:
    """Select the stored schedules and launch shape for an ensemble (exact member count and
    total interior points, else the nearest stored ensemble of the same build)."""
    from .runtime import KERNEL_RHS, KERNEL_RICCI

    p = ensemble_table_path(E)
    d = None
    if os.path.exists(p):
        with open(p) as f:
            d = json.load(f)
    elif os.path.isdir(TABLE_DIR):
        pre = f"ensemble_{build_tag(E.cfgs[0])}_"
        cands = sorted(fn for fn in os.listdir(TABLE_DIR) if fn.startswith(pre))
        if cands:
            with open(os.path.join(TABLE_DIR, cands[-1])) as f:
                d = json.load(f)
    if not d:
        return None
    if d["rhs"] in E.schedules(KERNEL_RHS):
        E.select_schedule(KERNEL_RHS, d["rhs"])
    if d["ricci"] in E.schedules(KERNEL_RICCI):
        E.select_schedule(KERNEL_RICCI, d["ricci"])
    E.set_launch(*d["launch"])
    return d


def tune_ensemble(E, steps: int = 3, write: bool = True, log=print) -> Dict:
    """Time every batched schedule pair, then every launch shape, on whole RK4 steps."""
    import torch

    from .runtime import KERNEL_RHS, KERNEL_RICCI

    saved = [st.y_n.clone() for st in E.states]
    stream = torch.cuda.current_stream()

    def run():
        for st, y in zip(E.states, saved):
            st.y_n.copy_(y)
        E.steps(steps)

    timings: Dict[str, float] = {}
    best = {}
    E.set_launch(64, 4, 0)
    for kind, label in ((KERNEL_RICCI, "ricci"), (KERNEL_RHS, "rhs")):
        res = {}
        for name in E.schedules(kind):
            E.select_schedule(kind, name)
            res[name] = _time(run, stream, 2)
            timings[f"{label}:{name}"] = res[name]
            log(f"ensemble {label} {name}: {res[name]:.3f} ms / {steps} steps")
        best[label] = min(res, key=res.get)
        E.select_schedule(kind, best[label])
    lres = {}
    for L in ENSEMBLE_LAUNCHES:
        E.set_launch(*L)
        lres[L] = _time(run, stream, 2)
        timings[f"launch:{L[0]}x{L[1]}:{L[2]}"] = lres[L]
        log(f"ensemble launch {L}: {lres[L]:.3f} ms / {steps} steps")
    launch = min(lres, key=lres.get)
    E.set_launch(*launch)
    for st, y in zip(E.states, saved):
        st.y_n.copy_(y)
    out = {"build": build_tag(E.cfgs[0]), "members": len(E.members), "rhs": best["rhs"], "ricci": best["ricci"],
           "launch": list(launch), "timings_ms": timings, "device": torch.cuda.get_device_name(0)}
    if write:
        os.makedirs(TABLE_DIR, exist_ok=True)
        with open(ensemble_table_path(E), "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
    return out


def main() -> None:
    import sys

    sys.path.insert(0, os.path.dirname(HERE))
    from bench import CONFIGS
    from .runtime import PatchConfig

    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c3")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=None, help="also copy the table here")
    a = ap.parse_args()
    if a.config == "c4":
        import numpy as np
        import torch

        from bench import c4_members
        from . import initial_data as ID
        from .runtime import Ensemble, PatchConfig

        mem = c4_members()
        E = Ensemble([PatchConfig(Nxx=m["Nxx"], grid_physical_size=10.0, eta=2.0) for m in mem])
        for P, st, m in zip(E.members, E.states, mem):
            y = ID.brill_lindquist(1, P.xx(), 10.0, 0.2, (m["m1"], m["m2"]), (0, 0, m["d"] / 2), (0, 0, -m["d"] / 2))
            st.y_n.copy_(torch.from_numpy(np.ascontiguousarray(y)).cuda())
            P.apply_bcs_outerextrap_and_inner(st.y_n)
        out = tune_ensemble(E)
        if a.out:
            os.makedirs(os.path.dirname(a.out), exist_ok=True)
            with open(a.out, "w") as f:
                json.dump(out, f, indent=1, sort_keys=True)
        print(json.dumps({k: out[k] for k in ("build", "members", "rhs", "ricci", "launch")}))
        return
    c = CONFIGS[a.config]
    cfg = PatchConfig(coord=c["coord"], Nxx=c["Nxx"], fd_order=c["fd_order"], symmetry_axes=c["sym"],
                      enable_KO=c["KO"], grid_physical_size=c["size"], SINHW=c["sinhw"], eta=2.0)
    out = tune(cfg, a.reps)
    if a.out:
        os.makedirs(os.path.dirname(a.out), exist_ok=True)
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps({k: out[k] for k in ("build", "Nxx", "rhs", "ricci", "tile")}))


if __name__ == "__main__":
    main()
