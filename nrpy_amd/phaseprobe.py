"""Where a workgroup-tiled stencil kernel spends a tile: per-phase timestamps.

A measurement tool, not part of the evolution path.  ``build`` takes one generated
workgroup-tiled rhs_eval schedule (``csrc/gen/bssn_<tag>_rhs_<schedule>.hip``) and writes an
instrumented copy: every wave reads ``s_memtime`` at the kernel start, before each phase's
box wait, after each phase's barrier, before the point algebra and at the end.  The stamps
are kept in LDS (a global store would enter the vector-memory counter the box waits count
on) and written out after the algebra, with the wave's hardware and XCD ids.  ``run`` launches the
copy once on the 192x96x96 SinhSpherical patch, checks its right-hand side against the
library's ``rhs_eval`` of the same schedule, and prints each share of a tile: prologue, waits
at the barriers, the phases (stencils formed from LDS), the point algebra, the mean per phase,
the spread of the four waves' arrival at a barrier and the gap between consecutive tiles on
a CU.

    python -m nrpy_amd.phaseprobe build [--schedule NAME]     (CPU: hipcc, no GPU needed)
    python -m nrpy_amd.phaseprobe run [--schedule NAME]       (GPU)
"""

from __future__ import annotations

import argparse
import ctypes
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
GEN = os.path.join(HERE, "csrc", "gen")
OUT = os.path.join(HERE, "lib", "phaseprobe")  # travels with the library; *.so is not committed
TAG = "SinhSpherical_symnone_KO1_fd6"
MAXMARKS = 62


def _default_schedule() -> str:
    import json

    with open(os.path.join(HERE, "autotune", f"{TAG}_192x96x96.json")) as f:
        return json.load(f)["rhs"]


def instrument(src: str) -> str:
    """The instrumented kernel source and a ``phaseprobe_launch`` entry point."""
    m = re.search(r"^(bssn_rhs_\w+_g0)\(", src, re.M)
    if not m or "BHAH_BARRIER_PHASE" not in src:
        raise ValueError("not a workgroup-tiled rhs_eval schedule")
    kname = m.group(1)
    src = src.replace("const double *__restrict__ rfm1, const bhah_kparams p) {",
                      "const double *__restrict__ rfm1, const bhah_kparams p, unsigned long long *__restrict__ ts) {", 1)
    sh = re.search(r"^  __shared__ double sh\[\d+\];\n", src, re.M)
    src = (src[:sh.end()] + "  __shared__ unsigned long long tsl[4][64];\n  int tsk_ = 0;\n"
           "  const unsigned hwid_ = __builtin_amdgcn_s_getreg((31 << 11) | 4);\n"
           "  const unsigned xcc_ = __builtin_amdgcn_s_getreg((31 << 11) | 20);\n  TSMARK();\n" + src[sh.end():])
    src = re.sub(r"(  BHAH_VMWAIT\(\d+\);\n)", r"  TSMARK();\n\1", src)
    src = re.sub(r"(  BHAH_BARRIER_PHASE\(\d+\);\n)", r"\1  TSMARK();\n", src)
    i = src.index("    if (i0 < p.Nx0 - p.NG && i1 < p.Nx1 - p.NG && i2 < p.Nx2 - p.NG) {")
    j = src.rindex("\n  {", 0, i)
    src = src[:j] + "\n  TSMARK();" + src[j:]
    k = src.index("\n}\n", i)
    src = src[:k] + """
  TSMARK();
  {
    const unsigned wv_ = threadIdx.x >> 6, ln_ = threadIdx.x & 63;
    const unsigned long long wid_ =
        (unsigned long long)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)) * 4 + wv_;
    if (ln_ < (unsigned)tsk_) ts[wid_ * 64 + ln_] = tsl[wv_][ln_];
    if (ln_ == 62) ts[wid_ * 64 + 62] = hwid_;
    if (ln_ == 63) ts[wid_ * 64 + 63] = ((unsigned long long)xcc_ << 32) | (unsigned)tsk_;
  }""" + src[k:]
    nmarks = src.count("TSMARK();")
    if nmarks > MAXMARKS:
        raise ValueError(f"{nmarks} marks: more than the {MAXMARKS} slots per wave")
    src = ("#define TSMARK() do { unsigned long long t_ = __builtin_amdgcn_s_memtime(); "
           "if ((threadIdx.x & 63) == 0) tsl[threadIdx.x >> 6][tsk_] = t_; tsk_++; } while (0)\n") + src
    m = re.search(r'extern "C" hipError_t launch_bssn_rhs_\w+\(.*?\n\}\n', src, re.S)
    src = src[:m.start()] + src[m.end():]
    return src + f"""
extern "C" int phaseprobe_launch(const double *in, const double *aux, double *out, const double *rfm0,
                                 const double *rfm1, const bhah_kparams *pp, unsigned long long *ts, int tx, int ty,
                                 int tz) {{
  const bhah_kparams p = *pp;
  hipLaunchKernelGGL({kname}, dim3((p.Nx0 - 2 * p.NG + tx - 1) / tx, (p.Nx1 - 2 * p.NG + ty - 1) / ty,
                     (p.Nx2 - 2 * p.NG + tz - 1) / tz), dim3(256), 0, 0, in, aux, out, rfm0, rfm1, p, ts);
  return (int)hipDeviceSynchronize();
}}
"""


def _paths(schedule: str):
    return (os.path.join(GEN, f"bssn_{TAG}_rhs_{schedule}.hip"), os.path.join(OUT, f"probe_{schedule}.hip"),
            os.path.join(OUT, f"probe_{schedule}.so"))


def build(schedule: str) -> str:
    from .build import HIPCC, _flags

    src, psrc, so = _paths(schedule)
    os.makedirs(OUT, exist_ok=True)
    with open(src) as f:
        text = instrument(f.read())
    with open(psrc, "w") as f:
        f.write(text)
    flags = [x for x in _flags(src)] + ["-shared", f"-I{GEN}"]
    r = subprocess.run([HIPCC, *flags, psrc, "-o", so], capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(r.stderr[-3000:])
    return so


class _KP(ctypes.Structure):
    _fields_ = [("Nx0", ctypes.c_int), ("Nx1", ctypes.c_int), ("Nx2", ctypes.c_int), ("NG", ctypes.c_int),
                ("Ntot", ctypes.c_long), ("s1", ctypes.c_long), ("s2", ctypes.c_long),
                ("invdxx0", ctypes.c_double), ("invdxx1", ctypes.c_double), ("invdxx2", ctypes.c_double),
                ("eta", ctypes.c_double), ("KO_gauge", ctypes.c_double), ("KO_nongauge", ctypes.c_double),
                ("wavespeed", ctypes.c_double), ("dt", ctypes.c_double), ("substep", ctypes.c_int)]


def run(schedule: str) -> dict:
    import numpy as np
    import torch

    sys.path.insert(0, os.path.dirname(HERE))
    from bench import CONFIGS

    from . import initial_data as ID
    from .codegen.hip_gen import _lds_geometry
    from .runtime import KERNEL_RHS, Patch, PatchConfig

    _, _, so = _paths(schedule)
    if not os.path.exists(so):
        raise RuntimeError(f"{so} is not built: run `python -m nrpy_amd.phaseprobe build` first")
    tx, ty, tz, _ = next(_lds_geometry(o) for o in schedule.split("_") if o[:1] == "L" and o[1:2].isdigit())
    c = CONFIGS["c3"]
    cfg = PatchConfig(coord=c["coord"], Nxx=c["Nxx"], fd_order=c["fd_order"], symmetry_axes=c["sym"],
                      enable_KO=c["KO"], grid_physical_size=c["size"], SINHW=c["sinhw"], eta=2.0,
                      KO_strength_gauge=0.3, KO_strength_nongauge=0.3)
    P = Patch(cfg, tuned=False)
    P.select_schedule(KERNEL_RHS, schedule)
    xx = P.xx()
    y = ID.brill_lindquist(c["coord"], xx, c["size"], c["sinhw"], c["masses"], c["pos1"], c["pos2"])
    y = torch.from_numpy(np.ascontiguousarray(y)).cuda()
    P.apply_bcs_outerextrap_and_inner(y)
    aux = P.new_gfs(6)
    P.Ricci_eval(y, aux)
    ref = P.new_gfs()
    P.rhs_eval(aux, y, ref)
    r0, r1 = (torch.from_numpy(np.ascontiguousarray(r)).cuda() for r in P.rfm())
    NG = c["fd_order"] // 2 + 1
    Nx = [n + 2 * NG for n in c["Nxx"]]
    kp = _KP(Nx[0], Nx[1], Nx[2], NG, Nx[0] * Nx[1] * Nx[2], Nx[0], Nx[0] * Nx[1],
             *[1.0 / (xx[d][1] - xx[d][0]) for d in range(3)], 2.0, 0.3, 0.3, 1.0, 0.0, 0)
    ntiles = ((c["Nxx"][0] + tx - 1) // tx) * ((c["Nxx"][1] + ty - 1) // ty) * ((c["Nxx"][2] + tz - 1) // tz)
    ts = torch.zeros(ntiles * 4 * 64, dtype=torch.int64, device="cuda")
    out = P.new_gfs()
    torch.cuda.synchronize()
    lib = ctypes.CDLL(so)
    lib.phaseprobe_launch.argtypes = [ctypes.c_void_p] * 5 + [ctypes.POINTER(_KP), ctypes.c_void_p] + [ctypes.c_int] * 3
    for _ in range(3):  # the last launch's stamps are kept
        rc = lib.phaseprobe_launch(y.data_ptr(), aux.data_ptr(), out.data_ptr(), r0.data_ptr(), r1.data_ptr(),
                                   ctypes.byref(kp), ts.data_ptr(), tx, ty, tz)
        if rc:
            raise RuntimeError(f"probe launch failed: hip error {rc}")
    scale = ref.abs().reshape(24, -1).max(dim=1).values.clamp_min(1e-300)
    rel = ((out - ref).abs().reshape(24, -1).max(dim=1).values / scale).max().item()
    return analyse(ts.cpu().numpy().reshape(-1, 64), rel)


def analyse(ts, rel_diff: float = float("nan")) -> dict:
    import numpy as np

    n = int(ts[0, 63] & 0xFFFFFFFF)
    if not ((ts[:, 63] & 0xFFFFFFFF) == n).all():
        raise RuntimeError("waves recorded different numbers of stamps")
    nph = (n - 3) // 2
    m = ts[:, :n].astype(np.float64)
    tile = m[:, n - 1] - m[:, 0]
    wait = np.stack([m[:, 2 + 2 * p] - m[:, 1 + 2 * p] for p in range(nph)], 1)
    comp = np.stack([m[:, 3 + 2 * p] - m[:, 2 + 2 * p] for p in range(nph)], 1)
    res = {
        "waves": int(len(ts)),
        "rel_diff_vs_rhs_eval": rel_diff,
        "tile_ticks": float(tile.mean()),
        "prologue": float((m[:, 1] - m[:, 0]).mean() / tile.mean()),
        "waits": float(wait.sum(1).mean() / tile.mean()),
        "phases": float(comp.sum(1).mean() / tile.mean()),
        "algebra": float((m[:, n - 1] - m[:, n - 2]).mean() / tile.mean()),
        "phase_wait_ticks": [round(float(v)) for v in wait.mean(0)],
        "phase_ticks": [round(float(v)) for v in comp.mean(0)],
    }
    arrive = m[:, 1:2 * nph:2].reshape(-1, 4, nph)
    res["barrier_arrival_spread_ticks"] = float((arrive.max(1) - arrive.min(1)).mean())
    hw = ts[:, 62].astype(np.int64)
    # CU within its XCD (cu_id, sh_id, se_id fields of HW_ID) and the XCD (XCC_ID)
    cu = ((hw >> 8) & 0xF) | (((hw >> 13) & 0x3) << 4) | (((hw >> 12) & 1) << 6) | ((ts[:, 63] >> 32) << 8)
    start, end = m[:, 0].reshape(-1, 4).min(1), m[:, n - 1].reshape(-1, 4).max(1)
    key = cu.reshape(-1, 4)[:, 0]
    gaps = []
    for k in np.unique(key):
        idx = np.where(key == k)[0]
        o = np.argsort(start[idx])
        gaps += list(start[idx][o][1:] - end[idx][o][:-1])
    res["gap_between_tiles_on_a_cu_ticks"] = float(np.median(gaps)) if gaps else None
    return res


def main() -> None:
    import json

    ap = argparse.ArgumentParser()
    ap.add_argument("action", choices=("build", "run"))
    ap.add_argument("--schedule", default=None, help="rhs_eval schedule (default: the tuned 192x96x96 one)")
    a = ap.parse_args()
    sched = a.schedule or _default_schedule()
    if a.action == "build":
        print(build(sched))
    else:
        print(json.dumps(run(sched)))


if __name__ == "__main__":
    main()
