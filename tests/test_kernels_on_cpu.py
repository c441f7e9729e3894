"""Generated rhs/Ricci schedules executed on the CPU.

Each schedule's translation unit is compiled as host C++ against a shim HIP
header (tests/kernel_cpu/) with AddressSanitizer, every array an exact-size
heap block, and run over the whole launch grid of a small patch; the outputs
must match the oracle to 1e-12 and any out-of-bounds access aborts the run.
This checks the emitted kernel text itself (stencil offsets, prefetch
placement, launch guards) without a GPU.
"""

import os
import re
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from nrpy_amd import initial_data as ID

HERE = os.path.dirname(os.path.abspath(__file__))
GEN = os.path.join(os.path.dirname(HERE), "nrpy_amd", "csrc", "gen")
SHIM = os.path.join(HERE, "kernel_cpu", "shim")
DRIVER = os.path.join(HERE, "kernel_cpu", "driver.cpp")

CASES = {
    "Spherical_sym2_KO0_fd4": dict(coord=1, Nxx=(12, 8, 2), fd_order=4, size=7.5, variant=0),
    "SinhSpherical_symnone_KO1_fd6": dict(coord=2, Nxx=(20, 10, 12), fd_order=6, size=64.0, variant=1),
    "SinhSpherical_symnone_KO1_fd4": dict(coord=2, Nxx=(14, 8, 10), fd_order=4, size=64.0, variant=1),
}


def schedules(tag):
    out = []
    for f in sorted(os.listdir(GEN)):
        m = re.match(rf"bssn_{tag}_(rhs|ricci)_(.+)\.hip$", f)
        if m:
            out.append((m.group(1), m.group(2)))
    return out


def build(tag, kind, cand, workdir):
    src = os.path.join(GEN, f"bssn_{tag}_{kind}_{cand}.hip")
    exe = os.path.join(workdir, f"{kind}_{cand}")
    # workgroup-tiled (LDS) schedules synchronise their threads: run each block's threads concurrently
    threaded = ["-DBHAH_CPU_THREADED=1", "-pthread"] if re.search(r"_L\d", cand) else []
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-w", *threaded,
           f"-I{SHIM}", f'-DKERNEL_FILE="{src}"', f"-DLAUNCH=launch_bssn_{kind}_{tag}_{cand}", DRIVER,
           "-o", exe, "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def run(exe, O, kind, y, aux, tile, workdir):
    inp, outp = os.path.join(workdir, "in.bin"), os.path.join(workdir, "out.bin")
    inv = [1.0 / (O.xx[d][1] - O.xx[d][0]) for d in range(3)]
    nout = 24 if kind == "rhs" else 6
    with open(inp, "wb") as f:
        np.array([*O.Nx, O.NG, tile[0], tile[1], 24, 6 if kind == "rhs" else 0, nout], dtype=np.int32).tofile(f)
        np.array(inv + [2.0, 0.3, 0.3, 1.0]).tofile(f)
        y.tofile(f)
        if kind == "rhs":
            aux.tofile(f)
        O.rfm0.tofile(f)
        O.rfm1.tofile(f)
    r = subprocess.run([exe, inp, outp], capture_output=True, text=True,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    assert r.returncode == 0, r.stderr[-3000:]
    return np.fromfile(outp).reshape(nout, *O.Nx[::-1])


def interior(a, NG):
    return a[:, NG:-NG, NG:-NG, NG:-NG]


def rel_err(a, b):
    num = np.abs(a - b).reshape(a.shape[0], -1).max(axis=1)
    den = np.abs(b).reshape(b.shape[0], -1).max(axis=1)
    return num / np.where(den > 0, den, 1.0)


def _check(oracle_lib, tag, cands, tiles, Nxx=None):
    c = dict(CASES[tag], **({"Nxx": Nxx} if Nxx else {}))
    O = oracle_lib.OracleGrid(c["coord"], c["Nxx"], c["fd_order"], c["size"], 0.2, variant=c["variant"])
    y = ID.brill_lindquist(O.coord, O.xx, O.g.p0, O.g.p1)
    y = np.ascontiguousarray(ID.smooth_perturbation(y, O.xx, 3))
    O.apply_bcs_outerextrap_and_inner(y)
    aux = O.ricci(y)
    ref = {"ricci": aux, "rhs": O.rhs(y, aux, O.params())}
    wd = tempfile.mkdtemp(prefix="kcpu_")
    try:
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            exes = list(ex.map(lambda kc: build(tag, kc[0], kc[1], wd), cands))
        for (kind, cand), exe in zip(cands, exes):
            for tile in tiles:
                out = run(exe, O, kind, y, aux, tile, wd)
                er = rel_err(interior(out, O.NG), interior(ref[kind], O.NG))
                assert er.max() < 1e-12, (kind, cand, tile, er)
    finally:
        shutil.rmtree(wd, ignore_errors=True)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_spherical_fd4_schedules_on_cpu(oracle_lib):
    _check(oracle_lib, "Spherical_sym2_KO0_fd4", schedules("Spherical_sym2_KO0_fd4")[:2], [(64, 4), (32, 8)])


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_blocked_tile_order_on_cpu(oracle_lib):
    """The blocked tile order (runs of 32 tiles as 2 x 4 x 4-tile blocks) only takes effect
    when the tile grid divides into blocks: 32x16x16 points are 2 x 4 x 4 tiles of 16x4x4, so
    every tile must be computed exactly once through the blocked mapping."""
    tag = "SinhSpherical_symnone_KO1_fd6"
    cands = [kc for kc in schedules(tag) if kc[1].endswith("_blk")]
    assert len(cands) == 2, cands
    _check(oracle_lib, tag, cands, [(64, 4)], Nxx=(32, 16, 16))


@pytest.mark.slow
@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("tag", list(CASES))
def test_every_schedule_on_cpu(oracle_lib, tag):
    _check(oracle_lib, tag, schedules(tag), [(64, 4), (32, 8)])


ENS_DRIVER = os.path.join(HERE, "kernel_cpu", "ens_driver.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("kind,cand,tile", [("ricci", "r1_lb1_pf16_u32", (64, 4, 1)),
                                            ("rhs", "p1_lb1_sl_lf_pf16_u32_w2", (64, 4, 1)),
                                            ("ricci", "r1_lb1_sj_L16x8x2x1_d3_dma_rb1", (16, 8, 2)),
                                            ("rhs", "p1_lb1_sj_L16x8x2x1_d3_dma_rb1", (16, 8, 2)),
                                            ("rhs", "p1_lb1_sj_L16x8x2x1_d3_dma_rb1_pg8", (16, 8, 2)),
                                            ("rhs", "p1_lb1_sj_L16x8x2x1_d3_dma_rb1_pg8_ub", (16, 8, 2))])
def test_ensemble_kernels_on_cpu(oracle_lib, kind, cand, tile):
    """Batched-ensemble form of a schedule over members of different grids, with a
    persistent grid smaller than the tile list, matches the oracle per member (LDS-tiled
    schedules walk tiles in their own units, each workgroup's threads run concurrently)."""
    tag = "Spherical_sym2_KO0_fd4"
    grids = [(12, 8, 2), (16, 6, 4), (10, 4, 2)]
    Os, ys, auxs, refs = [], [], [], []
    for k, N in enumerate(grids):
        O = oracle_lib.OracleGrid(1, N, 4, 5.0 + k, 0.2, variant=0)
        y = ID.brill_lindquist(O.coord, O.xx, O.g.p0, O.g.p1)
        y = np.ascontiguousarray(ID.smooth_perturbation(y, O.xx, 10 + k))
        O.apply_bcs_outerextrap_and_inner(y)
        aux = O.ricci(y)
        Os.append(O), ys.append(y), auxs.append(aux)
        refs.append(aux if kind == "ricci" else O.rhs(y, aux, O.params()))
    wd = tempfile.mkdtemp(prefix="kens_")
    try:
        src = os.path.join(GEN, f"bssn_{tag}_{kind}_{cand}.hip")
        exe = os.path.join(wd, "ens")
        threaded = ["-DBHAH_CPU_THREADED=1", "-pthread"] if tile[2] > 1 else []
        subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-w",
                        *threaded, f"-I{SHIM}", f'-DKERNEL_FILE="{src}"', f"-DLAUNCH=launch_bssn_ens_{kind}_{tag}_{cand}",
                        f"-DKIND_RHS={int(kind == 'rhs')}", ENS_DRIVER, "-o", exe, "-lm"],
                       check=True, capture_output=True)
        inp, outp = os.path.join(wd, "in.bin"), os.path.join(wd, "out.bin")
        with open(inp, "wb") as f:
            np.array([len(grids), tile[0], tile[1], 3, tile[2]], dtype=np.int32).tofile(f)
            for O, y, aux in zip(Os, ys, auxs):
                np.array([*O.Nx, O.NG], dtype=np.int32).tofile(f)
                np.array([1.0 / (O.xx[d][1] - O.xx[d][0]) for d in range(3)] + [2.0, 0.3, 0.3, 1.0]).tofile(f)
                y.tofile(f)
                if kind == "rhs":
                    aux.tofile(f)
                O.rfm0.tofile(f)
                O.rfm1.tofile(f)
        r = subprocess.run([exe, inp, outp], capture_output=True, text=True,
                           env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
        assert r.returncode == 0, r.stderr[-3000:]
        flat = np.fromfile(outp)
        nout = 24 if kind == "rhs" else 6
        off = 0
        for O, ref in zip(Os, refs):
            n = nout * O.Ntot
            out = flat[off:off + n].reshape(nout, *O.Nx[::-1])
            off += n
            er = rel_err(interior(out, O.NG), interior(ref, O.NG))
            assert er.max() < 1e-12, er
    finally:
        shutil.rmtree(wd, ignore_errors=True)
